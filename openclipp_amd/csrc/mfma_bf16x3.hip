// Kernel instances of the MFMA cross-correlation for one operand precision.
#define OCIP_MFMA_P 3
#include "mfma_inst.h"
