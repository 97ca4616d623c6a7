"""Synthetic inputs of the benchmark configurations (host side, numpy).

splitmix64 (Steele, Lea, Flood 2014; Vigna's constants) is counter based:
draw i (0-based) of seed s is mix(s + (i+1) * 0x9e3779b97f4a7c15), so host
and device generators agree draw for draw.  Every range is a power of two,
so a range is an exact right shift of the 64-bit draw:
  [0, 32767] = >> 49,  [0, 2^31) = >> 33,  [0, 15] = >> 60,  [0, 2^63) = >> 1,
  full dtype range = low bytes of the draw (shift 0, truncated).
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

SHIFT_15BIT = 49   # [0, 32767]
SHIFT_31BIT = 33   # [0, 2^31)
SHIFT_16VAL = 60   # [0, 15]
SHIFT_63BIT = 1    # [0, 2^63)

# per-configuration seeds
SEED_C1 = 0x5EED0001
SEED_C2_UNIFORM = 0x5EED0002
SEED_C2_FEW = 0x5EED0003
SEED_C3 = 0x5EED0004
SEED_C4_LENGTHS = 0x5EED0005
SEED_C4_KEYS = 0x5EED0006
SEED_C5 = 0x5EED0007


def splitmix64(seed, n, start=0):
    """draws start .. start+n-1 of the splitmix64 stream seeded with seed"""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed & (2**64 - 1)) + i * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def gen_keys(dtype, n, seed, shift):
    """keys[i] = (draw i >> shift) truncated to dtype (numpy array)"""
    dt = np.dtype(dtype)
    v = splitmix64(seed, n) >> np.uint64(shift)
    return v.astype(np.dtype(f"u{dt.itemsize}")).view(dt)


def unit_uniform(seed, n, start=0):
    """53-bit uniforms in [0, 1)"""
    return (splitmix64(seed, n, start) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def loguniform_lengths(seed, n, lo, hi):
    """floor(exp(ln lo + u (ln hi - ln lo))), clamped to [lo, hi]"""
    u = unit_uniform(seed, n)
    a, b = np.log(float(lo)), np.log(float(hi))
    v = np.floor(np.exp(a + u * (b - a))).astype(np.int64)
    return np.clip(v, lo, hi)


def segment_offsets(lengths):
    off = np.zeros(len(lengths) + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    return off


# configuration shapes
C1_N = 1_048_576
C2_N = 67_108_864
C3_ARRAYS, C3_SIZE = 65_536, 1024
C4_SEGMENTS, C4_LO, C4_HI = 20_000, 16, 262_144
C5_RATE, C5_SECONDS, C5_LO, C5_HI = 2000.0, 60.0, 1024, 4_194_304
C5_POOL_BYTES = 64 * 2**30
