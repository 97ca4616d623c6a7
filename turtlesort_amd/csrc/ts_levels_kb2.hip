// ts_levels_kb2.hip -- engine kernels for 2-byte keys (one file per key
// width so the instantiations compile in parallel).
#include "ts_levels_impl.h"

namespace ts {
levels_fn pick_kb2(bool idx, int S) { return pick_kb<2>(idx, S); }
}  // namespace ts
