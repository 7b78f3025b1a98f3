"""MI355X-native Stokes hot path (anisotropic P2-P1 assembly + Schur PCG)."""
