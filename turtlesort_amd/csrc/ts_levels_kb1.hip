// ts_levels_kb1.hip -- engine kernels for 1-byte keys (one file per key
// width so the instantiations compile in parallel).
#include "ts_levels_impl.h"

namespace ts {
levels_fn pick_kb1(bool idx, int S) { return pick_kb<1>(idx, S); }
}  // namespace ts
