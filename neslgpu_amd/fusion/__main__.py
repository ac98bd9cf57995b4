"""python -m neslgpu_amd.fusion defs.superops -o libfused.so [--hip out.hip]"""
import argparse

from .codegen import build


def main():
    p = argparse.ArgumentParser()
    p.add_argument("defs")
    p.add_argument("-o", "--output", required=True)
    p.add_argument("--hip", default=None, help="where to keep the generated HIP source")
    a = p.parse_args()
    with open(a.defs) as f:
        build(f.read(), a.output, a.hip, verbose=True)


if __name__ == "__main__":
    main()
