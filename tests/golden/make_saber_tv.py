"""Extract the Saber test vectors held by the reference into a JSON fixture.

Source: benlwk/DPCrypto SABER-GPU/include/tv.h:4-12 (GPL-3.0; test-vector
data only, no code).  The header is read as text: each ``name[...] = {...};``
array becomes a hex string (uint8 arrays) or a list of integers (uint16
arrays).  Run in a container where the reference is mounted:

    python tests/golden/make_saber_tv.py /root/reference/SABER-GPU/include/tv.h

Output: tests/golden/saber_tv.json
"""
import json
import os
import re
import sys

ARRAYS = {  # name -> element type in the header
    "A": "u16", "sp": "u16", "pk": "u8", "m_tv": "u16", "sk_tv": "u8", "buf_tv": "u8",
}


def main(path: str) -> None:
    text = open(path, encoding="ascii").read()
    out = {"source": "benlwk/DPCrypto SABER-GPU/include/tv.h:4-12"}
    for name, kind in ARRAYS.items():
        m = re.search(r"\b%s\s*\[[^\]]*\]\s*=\s*\{([^}]*)\}" % re.escape(name), text)
        if not m:
            raise SystemExit(f"array {name} not found")
        vals = [int(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()]
        if kind == "u8":
            out[name] = bytes(vals).hex()
        else:
            out[name] = vals
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "saber_tv.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print(dst, {k: (len(v) // 2 if isinstance(v, str) else len(v)) for k, v in out.items() if k != "source"})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/SABER-GPU/include/tv.h")
