// placeholder, filled below
#include "kernels.h"
