// FP32 instantiation of the kernels (FP32 sweeps, FP64 guards).
#include "tlq_launch_impl.hpp"

template struct tlq::Launch<float>;
