"""MI355X-native explicit residual + SSP-RK3 path of UCNS3D (WENO/MUSCL + HLLC)."""
