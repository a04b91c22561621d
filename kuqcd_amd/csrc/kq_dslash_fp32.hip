// Hopping-term instantiations: FP32 links (18-real long links and 12-real LinkF12).
#include "kq_dslash_impl.h"

namespace kq {

template void launch_dslash<LinkF, LinkF, float>(DslashArgs<LinkF, LinkF, float>, int, int, int, hipStream_t, int);
template void launch_dslash<LinkF, LinkF12, float>(DslashArgs<LinkF, LinkF12, float>, int, int, int, hipStream_t, int);

}  // namespace kq
