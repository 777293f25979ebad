// G1 MSM (filled in by the Pippenger engine).
#include "runtime.hpp"
