"""Reference-compatible host interface (``eng/`` of tiSPHi) over the HIP engine."""
