// Persistent findMode kernel instantiations for one model.
#include "resident_launch.h"

CCD_INSTANTIATE_LAUNCH(M_CLR)
