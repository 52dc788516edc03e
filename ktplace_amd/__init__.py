"""MI355X-native SimPL inner loop (B2B assembly + Jacobi-PCG + look-ahead
legalization) behind a C ABI; see include/ktplace_amd.h."""
