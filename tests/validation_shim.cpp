// Exposes the header-only argument checks of openclipp_amd/csrc/validation.h to the CPU
// tests (compiled with g++ by tests/test_abi.py; no HIP needed).
#include "validation.h"

extern "C" {
int vc_create_image(SImage i) { return ocip::check::create_image(i); }
int vc_proximity(SImage s, SImage t, SImage d) { return ocip::check::proximity(s, t, d); }
int vc_integral(SImage s, SImage d) { return ocip::check::integral(s, d); }
int vc_precision(SImage s, int p) { return ocip::check::precision(s, p); }
}
