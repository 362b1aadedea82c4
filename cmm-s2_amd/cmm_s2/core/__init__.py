"""Reference-compatible API of the CMM engine (cmm_s2/core of seth-tl/CMM-S2)."""
