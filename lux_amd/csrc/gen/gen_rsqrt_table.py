#!/usr/bin/env python3
"""Writes rsqrt_table.h: the two 4096-entry approximation tables of the x86 `rsqrtss` instruction as an Intel CPU
computes them.

For a normal positive input the instruction's result is a function of the top 12 mantissa bits and of the parity of
the exponent, scaled by a power of two. This script compiles a few lines of C with gcc, runs `rsqrtss` on the 4096
inputs 1.0 + i / 4096 (exponent field 127: the odd table) and on 2 * (1.0 + i / 4096) (exponent field 128: the even
table) on the host CPU and writes the results. It must run on an Intel CPU (vendor GenuineIntel): other vendors' tables
differ. Usage: python3 gen_rsqrt_table.py [out.h]
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r"""
#include <immintrin.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
int main(void) {
    for (uint32_t e = 127; e <= 128; e++)
        for (uint32_t i = 0; i < 4096; i++) {
            uint32_t b = (e << 23) | (i << 11), r;
            float x, y;
            memcpy(&x, &b, 4);
            y = _mm_cvtss_f32(_mm_rsqrt_ss(_mm_set_ss(x)));
            memcpy(&r, &y, 4);
            printf("%u\n", r);
        }
    return 0;
}
"""


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(HERE), "rsqrt_table.h")
    vendor = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("vendor_id")), "")
    if vendor != "GenuineIntel":
        sys.exit(f"needs an Intel CPU (this one is {vendor!r})")
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "t.c"), os.path.join(d, "t")
        open(c, "w").write(SRC)
        subprocess.check_call(["gcc", "-O1", "-msse", c, "-o", exe])
        vals = [int(v) for v in subprocess.check_output([exe]).split()]
    assert len(vals) == 8192
    lines = ["// Generated by gen/gen_rsqrt_table.py on an Intel CPU: the rsqrtss result for the input with top 12 mantissa",
             "// bits i and exponent field 127 (odd table) or 128 (even table); the low 11 bits do not change the result.",
             "// Do not edit by hand.", "#pragma once", "#include <stdint.h>"]
    for name, part in (("kRsqrtOddTable", vals[:4096]), ("kRsqrtEvenTable", vals[4096:])):
        lines.append(f"static const uint32_t {name}[4096] = {{")
        for i in range(0, 4096, 8):
            lines.append("    " + ", ".join(f"0x{v:08x}u" for v in part[i:i + 8]) + ",")
        lines.append("};")
    open(out, "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
