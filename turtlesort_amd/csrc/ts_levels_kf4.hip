// ts_levels_kf4.hip -- engine kernels for IEEE binary32 keys (one file per key
// type so the instantiations compile in parallel).
#include "ts_levels_impl.h"

namespace ts {
levels_fn pick_kf4(bool idx, int S) { return pick_kb<4 | KT_FLOAT>(idx, S); }
}  // namespace ts
