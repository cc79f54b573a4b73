// Pass-kernel launchers, double precision.
#define VIE_REAL double
#include "vie_launch_impl.h"

namespace vie {
bool size_supported(int n)
{
    switch (n) {
#define CASE(v) case v:
        VIE_SIZES(CASE)
#undef CASE
        return true;
    default:
        return false;
    }
}
}  // namespace vie
