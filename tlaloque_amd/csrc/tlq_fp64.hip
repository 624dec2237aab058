// FP64 instantiation of the kernels.  Built with -ffp-contract=off so that
// every multiply and add rounds separately, as in the reference's own build.
#include "tlq_launch_impl.hpp"

template struct tlq::Launch<double>;
