"""Copy the IJ file fixtures of the reference's own I/O tests into
tests/golden/io/ (run in a container that has the reference tree).

Sources (data files of src/test/TEST_ij, used by io.jobs and fsai.jobs):
  data/poisson_10x10x10_np1/IJ.{A,b,x0}.<iXfY>.00000.bin  binary IJ format
  data/beam_tet_dof459_np1/{A,b}.IJ.00000                 text IJ format
The expected results of io.jobs (io.saved) are recorded in
tests/golden/reference_runs.json under "io_ds_pcg".
"""
import os
import shutil
import sys

SRC = "/root/reference/src/test/TEST_ij/data"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "io")

FILES = [
    ("poisson_10x10x10_np1", "IJ.A.i4f8.00000.bin"),
    ("poisson_10x10x10_np1", "IJ.A.i8f4.00000.bin"),
    ("poisson_10x10x10_np1", "IJ.b.i4f8.00000.bin"),
    ("poisson_10x10x10_np1", "IJ.b.i4f4.00000.bin"),
    ("poisson_10x10x10_np1", "IJ.x0.i4f8.00000.bin"),
    ("poisson_10x10x10_np1", "IJ.x0.i8f4.00000.bin"),
    ("beam_tet_dof459_np1", "A.IJ.00000"),
    ("beam_tet_dof459_np1", "b.IJ.00000"),
]


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} not found")
    os.makedirs(DST, exist_ok=True)
    for sub, name in FILES:
        prefix = "poisson." if sub.startswith("poisson") else "beam."
        shutil.copyfile(os.path.join(SRC, sub, name), os.path.join(DST, prefix + name))
        print("copied", sub, name)


if __name__ == "__main__":
    main()
