// asv_march.hip -- one quarter of the marching pass's instances: steps
// 4P+1 .. 4P+4 for P = ASV_MARCH_PART (set by the build, 0..3).

#include "asv_march.h"

#ifndef ASV_MARCH_PART
#error "ASV_MARCH_PART (0..3) selects the steps this unit instantiates"
#endif

namespace asv {

namespace {

template <int K>
MarchFn pick(int mc, int sl, bool ht, bool counted, int lt)
{
    if (lt != 0) return (sl == 2 && !ht && !counted) ? march_fn16<K>(mc, lt) : nullptr;
    if (ht) return counted ? march_fn<K, true, true>(mc, sl) : march_fn<K, true, false>(mc, sl);
    return counted ? march_fn<K, false, true>(mc, sl) : march_fn<K, false, false>(mc, sl);
}

}  // namespace

template <>
MarchFn march_kernel_part<ASV_MARCH_PART>(int k, int mc, int sl, bool ht, bool counted, int lt)
{
    constexpr int k0 = 4 * ASV_MARCH_PART;
    switch (k - k0) {
    case 1: return pick<k0 + 1>(mc, sl, ht, counted, lt);
    case 2: return pick<k0 + 2>(mc, sl, ht, counted, lt);
    case 3: return pick<k0 + 3>(mc, sl, ht, counted, lt);
    case 4: return pick<k0 + 4>(mc, sl, ht, counted, lt);
    default: return nullptr;
    }
}

}  // namespace asv
