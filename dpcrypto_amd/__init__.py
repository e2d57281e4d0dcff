"""MI355X-native batched FrodoKEM and Saber KEM engine (FrodoKEM-976-SHAKE and Saber l=3 first).

The compute path is the HIP library ``libdpkem.so`` (C ABI in
``include/dpkem.h``); this package is its thin Python binding.
"""
from .kem import (FRODO976, SABER, FRODO976_AES, LIGHTSABER, FIRESABER, FRODO640, FRODO1344, MEM_HOST, MEM_DEVICE, FORM_BATCHED, FORM_PLAIN, TUNE_SABER_POLY_IPW, TUNE_FRODO_PERSIST, TUNE_FRODO_THREADS, SIZES,  # noqa: F401
                  DpkemError, Engine, Plan, OP_KEYGEN, OP_ENC, OP_DEC, PLAN_SHARE_MATRIX, PLAN_NO_PIPELINE, PLAN_NO_GRAPH, PLAN_NO_MERGE, ROLL_OVERLAP, load_library, header_functions, LIB_PATH,
                  Server, REQ_DTYPE, COMP_DTYPE, make_requests, now_us)
