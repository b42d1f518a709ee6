// kp_search1.hip -- a second build of the list-free one-word search kernel
// (blocks of at most 64 vertices), in a translation unit of its own: it is
// built with -structurizecfg-skip-uniform-regions, which leaves wave-uniform
// branches unstructured (scalar branches without exec-mask bookkeeping).
// That makes the one-word frame loop about 14% faster on the Chung-Lu 75k
// window, where the one-word class is the primary class of the epoch; it
// slows the wider kernels by about 5%, which therefore keep the default
// build, and so does the one-word class beside a wider primary class.
#define KP_ONE_WORD_SEARCH_ONLY
#include "kp_kernels.h"

namespace kpd {
template __global__ void search_kernel<1, 2, false, 1>(SearchArgs);
} // namespace kpd
