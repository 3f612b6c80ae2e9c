/* Forwarding header: Timer (reference Timer.h) is declared in ThunderEgg/amd.h. */
#pragma once
#include "amd.h"
