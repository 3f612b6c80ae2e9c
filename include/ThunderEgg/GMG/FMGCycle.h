/* Forwarding header: the hot-path API lives in ThunderEgg/amd.h. */
#pragma once
#include "../amd.h"
