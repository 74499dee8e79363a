"""MI355X-native closed-shell CCSD(T) hot path of SCESS.

Host-side mirror of the reference's method API (src/Quantum/CC/CC.h:99-164)
over the C ABI of libscess_gpu.so (include/scess_gpu.h). Submodules:

  scess_amd.synthetic  seeded synthetic inputs (numpy, host)
  scess_amd.gpu        ctypes binding of the C ABI; raises if the library is absent
  scess_amd.triples    RCCSD_TriplesCorrection and the TriplesFunction forward/backward pair
"""
