// Hopping-term instantiations: FP32 fat links with 16-bit long links.
#include "kq_dslash_impl.h"

namespace kq {

template void launch_dslash<LinkF, LinkH, float>(DslashArgs<LinkF, LinkH, float>, int, int, int, hipStream_t, int);
template void launch_dslash<LinkF, LinkH12, float>(DslashArgs<LinkF, LinkH12, float>, int, int, int, hipStream_t, int);

}  // namespace kq
