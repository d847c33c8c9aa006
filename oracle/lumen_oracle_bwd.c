/* Backward oracle: filled in below. */
#include "lumen_oracle.h"
