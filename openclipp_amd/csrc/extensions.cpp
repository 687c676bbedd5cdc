// extensions.cpp -- placeholder entry points for the batch and backward extensions.
#include "ocip_internal.h"

extern "C" {
ocipError ocip_API ocipCrossCorr_NormBatch(ocipImage, const ocipImage *, ocipImage *, uint, int)
{ return CL_INVALID_OPERATION; }
ocipError ocip_API ocipSqrDistance_NormBatch(ocipImage, const ocipImage *, ocipImage *, uint, int)
{ return CL_INVALID_OPERATION; }
}
