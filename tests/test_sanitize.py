"""The CPU emulation of the GPU kernels' index code under AddressSanitizer and UBSan:
tests/native/emu_sanitize.cpp (which compiles tests/native/block_emu.cpp and through it the product
headers) runs the forward decompositions (4- and 5-position suffix blocks, group order) against each
other, the descent's child indices at n=10, m=20, the suffix-tuple words, the sampling kernels'
index logic at the GPU tests' sizes (up to the 30,045,015 entries of a lossy n=10, m=20
distribution) and the device unrank of every layer of n=10, m=20, with every buffer access
bounds-checked. Built with g++ (host code only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_emulation_under_address_and_undefined_sanitizers(tmp_path):
    exe = tmp_path / "emu_sanitize"
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-Wno-unknown-pragmas", "-o", str(exe),
                    os.path.join(ROOT, "tests", "native", "emu_sanitize.cpp")], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "all checks passed" in r.stdout
