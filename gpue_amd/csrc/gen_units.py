"""Writes the per-(precision, N, plan) translation units of the sweep kernels.

Each unit instantiates one GPUE_SWEEP_TU (sweep_launch.h) so that the build of
the many kernel instances runs in parallel.  Run after changing SIZES; the
dispatch tables in sweep_dispatch.hip and plan_radices in gpue_ctx.hip list the
same sizes."""
import os

SIZES = [16, 32, 48, 64, 96, 128, 192, 256, 384, 512, 1024, 2048, 4096]

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    for n in SIZES:
        for prec, T in (("f64", "double"), ("f32", "float")):
            for v in (0, 1):
                name = f"sweep_{prec}_{n}_v{v}"
                text = (f"// written by gen_units.py: sweep kernels for N={n}, plan {v}, {T}\n"
                        f"#include \"sweep_launch.h\"\n"
                        f"GPUE_SWEEP_TU({T}, {n}, {v}, {name})\n")
                path = os.path.join(HERE, name + ".hip")
                old = open(path).read() if os.path.exists(path) else None
                if old != text:
                    with open(path, "w") as f:
                        f.write(text)


if __name__ == "__main__":
    main()
