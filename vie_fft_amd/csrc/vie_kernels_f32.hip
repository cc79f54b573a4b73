// Pass-kernel launchers, single precision.
#define VIE_REAL float
#include "vie_launch_impl.h"
