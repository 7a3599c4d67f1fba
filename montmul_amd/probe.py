"""Issue-rate probe of the integer instructions the step kernel is built from (mm_probe_issue).

    python -m montmul_amd.probe [--iters 20000] [--blocks 0] [--ops 0,1,2,3,4]

Each op runs, per thread, 8 independent chains of one instruction for `iters` iterations, over
`blocks` workgroups of 256 threads (0: CUs x 32, which is 128 waves per CU over the launch). Prints one JSON line
per op: device time, chained instructions per second (waves x 8 x iters / time) and that rate
per CU per cycle at the clock given by --clock-ghz (the counter run under rocprofv3 gives the
measured clock and SQ_INSTS_VALU). The step kernel's roofline divides its own instruction rate
by the rate measured here.
"""

import argparse
import ctypes as C
import json

import montmul_amd as mm

OPS = {0: "v_mad_u64_u32", 1: "v_add_u32", 2: "v_mul_hi_u32", 3: "v_lshl_add_u64", 4: "v_mad_u64_u32 + v_add_u32"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--iters", type=int, default=20000)
    ap.add_argument("--blocks", type=int, default=0)
    ap.add_argument("--ops", default="0,1,2,3,4")
    ap.add_argument("--clock-ghz", type=float, default=2.4)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args()
    hip = mm.Hip(device=args.device)
    cus = hip.cu_count()
    hip.close()
    blocks = args.blocks or cus * 32
    for op in [int(x) for x in args.ops.split(",")]:
        for rep in range(2):  # the first launch also loads the code
            ms, ck = C.c_double(), C.c_uint64()
            if mm.lib.mm_probe_issue(args.device, op, blocks, args.iters, C.byref(ms), C.byref(ck)) != 0:
                raise SystemExit((mm.lib.mm_probe_last_error() or b"").decode())
        waves = blocks * 256 // 64
        per_iter = 16 if op == 4 else 8
        instr = waves * per_iter * args.iters
        rate = instr / (ms.value * 1e-3)
        print(json.dumps({"op": op, "instr": OPS[op], "blocks": blocks, "iters": args.iters, "ms": round(ms.value, 3),
                          "wave_instr_per_s": rate, "per_cu_per_cycle": rate / (cus * args.clock_ghz * 1e9),
                          "cus": cus, "clock_ghz_assumed": args.clock_ghz, "checksum": ck.value}), flush=True)


if __name__ == "__main__":
    main()
