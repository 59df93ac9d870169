"""MI355X-native stochastic-DFT Chebyshev filter path (host side).

Submodules:
  systems : synthetic systems (cells, plane-wave sets, FFTW maps, v_KS, KB tables, noise)
  hip     : ctypes binding of libsdft_hip.so (the C ABI in include/sdft_hip.h)
  stodft  : the reference's filter-path operations over the C ABI
  build   : in-tree build of the HIP and host libraries
  pinyfiles : the reference's checkpoint, density and fragment-orbital text files
"""
