// stlp_lbfgs: placeholder until the L-BFGS port lands.
#include <string>
#include "../../include/stl_bridge.h"
extern "C" int64_t stlp_lbfgs(double*, size_t, const char*, stlp_callback, void*) { return -1; }
