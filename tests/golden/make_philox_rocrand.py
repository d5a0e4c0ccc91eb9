"""Write philox_rocrand.json: Philox4x32-10 outputs from rocRAND's open-source
header (rocrand_device::philox4x32_10_engine::ten_rounds, compiled host-side
with hipcc) for seeded random counters and keys.  An implementation
independent of the oracle's, used to pin it.

Run from the repository root:  python tests/golden/make_philox_rocrand.py
"""
import json
import os
import subprocess
import tempfile

import numpy as np

SRC = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;  // protected in the header
};
int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    Engine e;
    while (scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 c = {c0, c1, c2, c3};
        uint2 k = {k0, k1};
        uint4 o = e.ten_rounds(c, k);
        printf("%u %u %u %u\n", o.x, o.y, o.z, o.w);
    }
    return 0;
}
"""


def main():
    rng = np.random.default_rng(20240601)
    inp = rng.integers(0, 2**32, size=(64, 6), dtype=np.uint64)
    inp[0] = 0
    inp[1] = 2**32 - 1
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "p.cpp")
        exe = os.path.join(d, "p")
        open(src, "w").write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-o", exe, src], check=True)
        txt = "\n".join(" ".join(str(int(v)) for v in row) for row in inp) + "\n"
        out = subprocess.run([exe], input=txt, capture_output=True, text=True, check=True).stdout
    rows = [list(map(int, l.split())) for l in out.strip().splitlines()]
    cases = [{"ctr": [int(v) for v in inp[i, :4]], "key": [int(v) for v in inp[i, 4:]], "out": rows[i]}
             for i in range(len(rows))]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "philox_rocrand.json")
    json.dump({"source": "rocrand_philox4x32_10.h ten_rounds (ROCm 7.2)", "cases": cases}, open(path, "w"), indent=0)
    print(f"wrote {len(cases)} cases to {path}")


if __name__ == "__main__":
    main()
