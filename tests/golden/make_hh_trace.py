"""Extract the HH golden voltage trace (201 samples, 0.1 ms apart, rounded to 0.01 mV) from the
reference's integration test neuwon/rxd/tests/test_model.py:77-113 into a plain data file.

Reads the reference test as text; nothing from the reference is imported or executed.
Usage: python tests/golden/make_hh_trace.py /root/reference
"""
import os
import re
import sys


def main(ref_root):
    src = open(os.path.join(ref_root, "neuwon/rxd/tests/test_model.py")).read()
    m = re.search(r"correct = TimeSeries\(\s*\[(.*?)\]\s*,", src, re.S)
    values = [float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()]
    assert len(values) == 201
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hh_trace_test_model.txt")
    with open(out, "w") as f:
        f.write("# tip voltage (mV) every 0.1 ms, t = 0..20 ms; neuwon/rxd/tests/test_model.py:77-113\n")
        f.write("\n".join(f"{v:.2f}" for v in values) + "\n")
    print(out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
