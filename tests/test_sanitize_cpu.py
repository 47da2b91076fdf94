"""Host C++ under AddressSanitizer and UndefinedBehaviorSanitizer.

The voxelizer of libwt (wind-tunnel_amd/csrc/wt_voxel.cpp: ray parity, shell
raster, the run-based flood-fill closure, analytic ellipsoid bodies) and the
CPU oracle (oracle/wt_oracle.cpp: every storage form's step, force, link
classes, particles, slices, Cp) are compiled by g++ with
-fsanitize=address,undefined (no recovery) into one program with
tests/sanitize_main.cpp, which runs the eight voxelizer golden cases (masks
checked bit for bit) and small flows through the oracle. Any out-of-bounds
access, use after free, leak, overflow or other undefined operation fails the
run with a sanitizer report."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from test_lib_cpu import _golden_cases

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which('g++') is None, reason='needs g++')
def test_voxelizer_and_oracle_are_sanitizer_clean():
    tmp = tempfile.mkdtemp(prefix='wt_san_')
    try:
        exe = os.path.join(tmp, 'sanitize_main')
        cmd = ['g++', '-std=c++17', '-O1', '-g', '-fno-omit-frame-pointer', '-fopenmp',
               '-fsanitize=address,undefined,float-cast-overflow', '-fno-sanitize-recover=all',
               '-ffp-contract=off', '-I', os.path.join(REPO, 'include'),
               '-I', os.path.join(REPO, 'wind-tunnel_amd', 'csrc'),
               os.path.join(REPO, 'tests', 'sanitize_main.cpp'),
               os.path.join(REPO, 'wind-tunnel_amd', 'csrc', 'wt_voxel.cpp'),
               os.path.join(REPO, 'oracle', 'wt_oracle.cpp'), '-o', exe]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        assert r.returncode == 0, r.stdout
        cases = os.path.join(tmp, 'cases')
        os.makedirs(cases)
        with open(os.path.join(cases, 'index.txt'), 'w') as idx:
            for c, tris, mask in _golden_cases():
                nx, ny, nz = c['dims']
                tris.astype(np.float32).tofile(os.path.join(cases, c['name'] + '.tris'))
                mask.astype(np.uint8).tofile(os.path.join(cases, c['name'] + '.mask'))
                idx.write(f"{c['name']} {nx} {ny} {nz} {int(c['ground'])} {c['ntris']} {c['count']} "
                          f"{c['frontal']} {c['planform']}\n")
        env = dict(os.environ, ASAN_OPTIONS='detect_leaks=1:abort_on_error=0:halt_on_error=1',
                   UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1', OMP_NUM_THREADS='2')
        r = subprocess.run([exe, cases], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                           timeout=600)
        assert r.returncode == 0 and 'sanitize ok' in r.stdout, r.stdout[-4000:]
        assert 'runtime error' not in r.stdout and 'AddressSanitizer' not in r.stdout, r.stdout[-4000:]
        assert r.stdout.count('voxelize ') == 8
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
