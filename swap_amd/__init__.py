"""MI355X-native hybrid-PIC time step of SWAP (gfx950 HIP kernels behind a C-ABI)."""
from .api import (B1, B12, B1P2, BTC, NP, UP, AJ, E, NU, AJC, UPC, TMP0, TMP1, TMP2,  # noqa
                  FP32, FP64, SPECIES, TRACK_IDS, Swap, SwapError, Cloud, load_library,
                  LIB_PATH)
from . import physics  # noqa: F401
