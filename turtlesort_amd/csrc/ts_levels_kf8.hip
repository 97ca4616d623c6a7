// ts_levels_kf8.hip -- engine kernels for IEEE binary64 keys (one file per key
// type so the instantiations compile in parallel).
#include "ts_levels_impl.h"

namespace ts {
levels_fn pick_kf8(bool idx, int S) { return pick_kb<8 | KT_FLOAT>(idx, S); }
}  // namespace ts
