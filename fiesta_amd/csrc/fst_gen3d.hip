// gen3d's fused advection (launch_gen_adv: k_adv3d_lds without the cartesian
// pressure-gradient epilogue) in a translation unit of its own.  Instantiated
// beside the cartesian stage kernels, its variants changed the code generated
// for the FP64 cartesian advection kernel (about 100 more instructions, 2 %
// slower); compiled apart, that kernel's code is what it was.
#define FST_GEN3D_TU
#include "fst_fused3d.hip"
