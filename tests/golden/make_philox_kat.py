"""Writes philox_kat.json: Philox4x32-10 blocks computed by rocRAND's
philox4x32_10 engine (an implementation independent of this repository),
compiled for the host from /opt/rocm/include/rocrand/rocrand_philox4x32_10.h.

Run in a container with ROCm: python tests/golden/make_philox_kat.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

# counters/keys: the Random123 known-answer inputs plus walk-engine-shaped ones
CASES = [
    ([0, 0, 0, 0], [0, 0]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF, 0xFFFFFFFF]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]),
    ([1, 0, 0, 0], [1, 0]),
    ([12345, 0, 7, 0], [1, 0]),
    ([99, 1, 0xFFFFFFFF, 3], [0xDEADBEEF, 0x01234567]),
    ([0x524D4154, 0, 2, 0x524D4154], [7, 0]),
]

SRC = r"""
#include <cstdio>
#include <rocrand/rocrand_philox4x32_10.h>
struct Engine : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;  // protected in rocRAND
};
int main() {
    unsigned int c[4], k[2];
    while (scanf("%u %u %u %u %u %u", &c[0], &c[1], &c[2], &c[3], &k[0], &k[1]) == 6) {
        Engine e;
        uint4 r = e.ten_rounds(uint4{c[0], c[1], c[2], c[3]}, uint2{k[0], k[1]});
        printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
"""


def main():
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "kat.cpp")
        exe = os.path.join(td, "kat")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-o", exe, src], check=True)
        inp = "".join(" ".join(str(x) for x in c + k) + "\n" for c, k in CASES)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    vec = []
    for (c, k), line in zip(CASES, out):
        vec.append({"ctr": c, "key": k, "out": [int(x) for x in line.split()]})
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump({"generator": "rocRAND philox4x32_10_engine::ten_rounds (host build)", "vectors": vec}, f,
                  indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
