// ts_levels_kb4.hip -- engine kernels for 4-byte keys (one file per key
// width so the instantiations compile in parallel).
#include "ts_levels_impl.h"

namespace ts {
levels_fn pick_kb4(bool idx, int S) { return pick_kb<4>(idx, S); }
}  // namespace ts
