"""Build of the native library and of the oracle used by the tests."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(directory, jobs=8):
    subprocess.run(["make", f"-j{jobs}", "-C", directory], check=True)


def build_native(jobs=8, force=True):
    """Compile symboliclib_amd/lib/libsymlib.so (always rebuilt when force)."""
    pkg = os.path.join(ROOT, "symboliclib_amd")
    if force:
        subprocess.run(["make", "-C", pkg, "clean"], check=True, stdout=subprocess.DEVNULL)
    _make(pkg, jobs)


def build_oracle(force=True):
    d = os.path.join(ROOT, "oracle")
    if force:
        subprocess.run(["make", "-C", d, "clean"], check=True, stdout=subprocess.DEVNULL)
    _make(d)
