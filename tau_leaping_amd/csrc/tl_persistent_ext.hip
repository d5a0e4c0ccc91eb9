// tl_persistent_ext.hip -- the persistent kernels again, for models with
// sources or channels of three reactant molecules (Sa+Sb+Sc, 2Sa+Sb), or runs
// with the exact Poisson sampler; the host picks them through kernel_for_ext.
#define TL_PERSISTENT_EXT 1
#include "tl_persistent.hip"
