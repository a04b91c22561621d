// Hopping-term instantiations: FP64 links.
#include "kq_dslash_impl.h"

namespace kq {

template void launch_dslash<LinkD, LinkD, double>(DslashArgs<LinkD, LinkD, double>, int, int, int, hipStream_t, int);

}  // namespace kq
