/* placeholder: linear operators added next */
#include "mlref.h"
