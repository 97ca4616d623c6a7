// ts_levels_kb8.hip -- engine kernels for 8-byte keys (one file per key
// width so the instantiations compile in parallel).
#include "ts_levels_impl.h"

namespace ts {
levels_fn pick_kb8(bool idx, int S) { return pick_kb<8>(idx, S); }
}  // namespace ts
