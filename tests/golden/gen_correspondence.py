"""Write correspondence_M12.txt from TSEMLib's own multi-index numbering.

Runs oracle/_ref/corr_dump (compiled by `make -C oracle ref` from the
reference's include/Correspondence.h) for M = 12 and stores its output: the
(number, multi-index, index2number) table and the degree block bounds."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle", "_ref", "corr_dump")

if __name__ == "__main__":
    out = subprocess.check_output([EXE, "12"]).decode()
    with open(os.path.join(HERE, "correspondence_M12.txt"), "w") as f:
        f.write(out)
    print("wrote", len(out.splitlines()), "lines")
