"""CPU checks of the register-resident sweep's layout (tests/cpp/plan_check.cpp): phase layouts and the LDS
swizzle are bijections, the tuned plans are free of LDS bank conflicts, and the phase schedule emulated on
the CPU reproduces the oracle's sweep bit for bit."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_plan_check():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True, capture_output=True)
    subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp")], check=True, capture_output=True)
    r = subprocess.run([os.path.join(ROOT, "tests", "cpp", "bin", "plan_check")], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
