"""The batch arenas' first-fit range allocator (tpm_alloc.hpp), host code:
randomized allocate / release sequences against a brute-force occupancy
model -- same offsets (first fit over coalesced free runs), no overlap,
the high-water mark is the end of the highest live range."""
import os
import random
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include "tpm_alloc.hpp"
int main() {
    tpm_range_alloc a;
    long long cap;
    if (std::scanf("%lld", &cap) != 1) return 1;
    a.reset(cap);
    char op;
    long long x, y;
    while (std::scanf(" %c", &op) == 1) {
        if (op == 'A') {
            std::scanf("%lld %lld", &x, &y);
            std::printf("%lld\n", (long long)a.alloc(x, y));
        } else if (op == 'R') {
            std::scanf("%lld %lld", &x, &y);
            a.release(x, y);
        } else if (op == 'H') {
            std::printf("%lld\n", (long long)a.high_water());
        } else if (op == 'F') {
            std::printf("%zu\n", a.free_.size());
        }
    }
    return 0;
}
"""


def model_alloc(used, size, align):
    cap = len(used)
    i = 0
    while i < cap:
        if used[i]:
            i += 1
            continue
        j = i
        while j < cap and not used[j]:
            j += 1
        off = (i + align - 1) // align * align
        if off + size <= j:
            for k in range(off, off + size):
                used[k] = True
            return off
        i = j
    return -1


def test_first_fit_allocator_matches_model(tmp_path):
    src = tmp_path / "alloc.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "alloc"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "tigerpm_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    rng = random.Random(11)
    cap = 4096
    used = [False] * cap
    live = []
    ops, expect = [str(cap)], []
    for _ in range(3000):
        if live and rng.random() < 0.45:
            off, size = live.pop(rng.randrange(len(live)))
            for k in range(off, off + size):
                used[k] = False
            ops.append("R %d %d" % (off, size))
        else:
            size, align = rng.randint(1, 300), rng.choice([1, 32, 64])
            off = model_alloc(used, size, align)
            if off >= 0:
                live.append((off, size))
            ops.append("A %d %d" % (size, align))
            expect.append(off)
        ops.append("H")
        hw = max((o + s for o, s in live), default=0)
        expect.append(hw)
    ops.append("F")
    runs = sum(1 for k in range(cap) if not used[k] and (k == 0 or used[k - 1]))
    expect.append(runs)
    out = subprocess.run([str(exe)], input="\n".join(ops) + "\n", capture_output=True, text=True, check=True)
    got = [int(v) for v in out.stdout.split()]
    assert got == expect
