"""The drop-in staging copy pool (csrc/copy_pool.hpp) under fork, on the
CPU: a parent that has used the pool forks; the child copies several
times and exits (running the pool's destructor), and the parent keeps
copying afterwards.  Every copy must be exact and nothing may hang."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = r'''
#include "copy_pool.hpp"
#include <sys/wait.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
static oxav1::CopyPool pool;
static int check(size_t n, int salt) {
    std::vector<unsigned char> a(n), b(n);
    for (size_t i = 0; i < n; i++) a[i] = (unsigned char)(i * 131 + salt);
    pool.copy(b.data(), a.data(), n);
    return a == b ? 0 : 1;
}
int main() {
    int bad = 0;
    for (int i = 0; i < 3; i++) bad |= check((3u << 20) + 17 * i, i);
    pid_t c = fork();
    if (c == 0) {
        int cb = 0;
        for (int i = 0; i < 4; i++) cb |= check((2u << 20) + 4099 * i, 7 + i);
        std::exit(cb ? 3 : 0);  // static destructors run here
    }
    int st = 0;
    waitpid(c, &st, 0);
    if (!WIFEXITED(st) || WEXITSTATUS(st) != 0) return 4;
    for (int i = 0; i < 3; i++) bad |= check((5u << 20) + 3 * i, 20 + i);
    return bad;
}
'''


def test_copy_pool_survives_fork(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no C++ compiler")
    src = tmp_path / "fork_pool.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "fork_pool"
    subprocess.run([gxx, "-O2", "-std=c++17", "-pthread",
                    f"-I{os.path.join(ROOT, 'oxideav-av1_amd', 'csrc')}", str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], timeout=60)
    assert r.returncode == 0
