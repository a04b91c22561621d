// Hopping-term instantiations: 16-bit fat and long links.
#include "kq_dslash_impl.h"

namespace kq {

template void launch_dslash<LinkH, LinkH, float>(DslashArgs<LinkH, LinkH, float>, int, int, int, hipStream_t, int);
template void launch_dslash<LinkH, LinkH12, float>(DslashArgs<LinkH, LinkH12, float>, int, int, int, hipStream_t, int);

}  // namespace kq
