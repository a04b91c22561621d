// Hopping-term instantiations: long links in 12 reals with a complex scale (U(3) products), FP32
// and 16-bit.
#include "kq_dslash_impl.h"

namespace kq {

template void launch_dslash<LinkF, LinkF12C, float>(DslashArgs<LinkF, LinkF12C, float>, int, int, int, hipStream_t,
                                                    int);
template void launch_dslash<LinkF, LinkH12C, float>(DslashArgs<LinkF, LinkH12C, float>, int, int, int, hipStream_t,
                                                    int);
template void launch_dslash<LinkH, LinkH12C, float>(DslashArgs<LinkH, LinkH12C, float>, int, int, int, hipStream_t,
                                                    int);

}  // namespace kq
