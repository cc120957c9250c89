"""MI355X-native FEMuS multigrid FEM hot path (HIP, gfx950).

The product is the C-ABI library libfemus_hip.so (include/femus_hip.h);
exfemus_amd.femus mirrors the reference's MGSolBase/MGSolT/MGSolNS interface
over it.  exfemus_amd.configs holds the benchmark configurations.
"""
