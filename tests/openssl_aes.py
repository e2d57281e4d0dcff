"""AES-ECB from the system's OpenSSL libcrypto through ctypes, as an
independent implementation to check ours against (tests only)."""
import ctypes
import ctypes.util


def aes_ecb():
    """enc(key, data) -> bytes for 16- or 32-byte keys, or None without libcrypto."""
    path = ctypes.util.find_library("crypto")
    if not path:
        return None
    lib = ctypes.CDLL(path)
    lib.EVP_CIPHER_CTX_new.restype = ctypes.c_void_p
    lib.EVP_aes_128_ecb.restype = ctypes.c_void_p
    lib.EVP_aes_256_ecb.restype = ctypes.c_void_p
    lib.EVP_EncryptInit_ex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                       ctypes.c_char_p]
    lib.EVP_CIPHER_CTX_set_padding.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.EVP_EncryptUpdate.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                      ctypes.c_char_p, ctypes.c_int]
    lib.EVP_CIPHER_CTX_free.argtypes = [ctypes.c_void_p]

    def enc(key, data):
        ctx = lib.EVP_CIPHER_CTX_new()
        cipher = lib.EVP_aes_128_ecb() if len(key) == 16 else lib.EVP_aes_256_ecb()
        assert lib.EVP_EncryptInit_ex(ctx, cipher, None, key, None) == 1
        lib.EVP_CIPHER_CTX_set_padding(ctx, 0)
        out = ctypes.create_string_buffer(len(data) + 32)
        n = ctypes.c_int(0)
        assert lib.EVP_EncryptUpdate(ctx, out, ctypes.byref(n), data, len(data)) == 1
        lib.EVP_CIPHER_CTX_free(ctx)
        return out.raw[: n.value]
    return enc
