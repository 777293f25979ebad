/* quotient numerator oracle: filled in below */
#include "oracle.h"
