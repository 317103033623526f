#!/usr/bin/env python3
"""Writes rcp_table.h: the 4096-entry approximation table of the x86 `rcpss` instruction as an Intel CPU computes it.

The instruction's result for a normal input is a function of the top 12 mantissa bits (the table entry for an input
with exponent field 127) scaled by the input's power of two. This script compiles a few lines of C with gcc, runs
`rcpss` on the 4096 inputs 1.0 + i / 4096 on the host CPU and writes the results. It must run on an Intel CPU
(vendor GenuineIntel): other vendors' tables differ. Usage: python3 gen_rcp_table.py [out.h]
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
    for (uint32_t i = 0; i < 4096; i++) {
        uint32_t b = 0x3f800000u | (i << 11), r;
        float x, y;
        memcpy(&x, &b, 4);
        y = _mm_cvtss_f32(_mm_rcp_ss(_mm_set_ss(x)));
        memcpy(&r, &y, 4);
        printf("%u\n", r);
    }
    return 0;
}
"""


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(HERE), "rcp_table.h")
    vendor = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("vendor_id")), "")
    if vendor != "GenuineIntel":
        sys.exit(f"needs an Intel CPU (this one is {vendor!r})")
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "t.c"), os.path.join(d, "t")
        open(c, "w").write(SRC)
        subprocess.check_call(["gcc", "-O1", "-msse", c, "-o", exe])
        vals = [int(v) for v in subprocess.check_output([exe]).split()]
    assert len(vals) == 4096
    lines = ["// Generated by gen/gen_rcp_table.py on an Intel CPU: the rcpss result for the input with exponent field 127",
             "// and top 12 mantissa bits i (the low 11 bits do not change the result). Do not edit by hand.",
             "#pragma once", "#include <stdint.h>", "#define LUX_RCP_TABLE_INIT { \\"]
    for i in range(0, 4096, 8):
        lines.append("    " + ", ".join(f"0x{v:08x}u" for v in vals[i:i + 8]) + ", \\")
    lines.append("}")
    lines.append("static const uint32_t kRcpTable[4096] = LUX_RCP_TABLE_INIT;")
    open(out, "w").write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
