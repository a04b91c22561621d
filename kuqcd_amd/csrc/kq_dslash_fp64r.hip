// Hopping-term instantiations: FP64 fat links with FP64 long links in 14 reals.
#include "kq_dslash_impl.h"

namespace kq {

template void launch_dslash<LinkD, LinkD14, double>(DslashArgs<LinkD, LinkD14, double>, int, int, int, hipStream_t,
                                                    int);

}  // namespace kq
