"""Host check of the wavefront kernel's digestion lane maps
(byteqc_amd/csrc/digest_split.h): lane-group sums equal the direct
four-index sums for every class up to (ff|ff), each element is produced by
exactly one lane and every block and density index stays in range."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_digest_lane_maps(tmp_path):
    gxx = shutil.which('g++')
    if gxx is None:
        pytest.skip('no host C++ compiler')
    exe = str(tmp_path / 'digest_split_check')
    subprocess.run([gxx, '-std=c++17', '-O1', '-o', exe,
                    os.path.join(HERE, 'cpp', 'digest_split_check.cpp')], check=True)
    out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip().endswith('OK'), out.stdout
