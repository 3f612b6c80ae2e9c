/* Umbrella header of the hot-path API (ThunderEgg/amd.h). */
#pragma once
#include "ThunderEgg/amd.h"
