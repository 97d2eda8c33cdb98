// shared_libm.cpp -- TEST INFRASTRUCTURE ONLY.  C entry points of the
// elementary functions the adjoint's backward uses for its own divided
// differences (trinity_amd/csrc/trn_libm.hpp), so that the CPU reference
// adjoint (trinity_adjoint.c) rounds them exactly as the device does: the
// gradient moves by about 3e-10 of its largest entry per ulp of them.  The
// reference's own functions (the forward energy) keep calling the host libm.
#include "../trinity_amd/csrc/trn_libm.hpp"

extern "C" {
double orc_shared_log(double x) { return trn::libm::log(x); }
double orc_shared_exp(double x) { return trn::libm::exp(x); }
double orc_shared_log1p_small(double r) { return trn::libm::log1p_small(r); }
double orc_shared_expm1(double d) { return trn::libm::expm1(d); }
}
