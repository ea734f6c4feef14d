"""Structure of the compiled persistent tile kernels (gfx950 assembly of this build; no GPU).

The persistent kernel claims a tile with one thread, then every wave passes two
barriers and reads the claim.  If the compiler threads a divergent block-leader
branch across the loop back-edge, the barriers after the claim end up in a
nested loop that the other lanes re-enter without the claim, and a workgroup
reprocesses one tile forever.  Every variant must keep the claim and the two
barriers at the same loop level.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


@pytest.fixture(scope='module')
def device_asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip('hipcc not available')
    out = tmp_path_factory.mktemp('asm') / 'rqx.s'
    cmd = [HIPCC, '--offload-arch=gfx950', '-O3', '-std=c++17', '-Xclang', '-target-feature', '-Xclang',
           '-packed-fp32-ops', '-I', os.path.join(ROOT, 'include'), '--cuda-device-only', '-S',
           os.path.join(ROOT, 'rqutils_amd', 'csrc', 'rqx.hip'), '-o', str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text().split('\n')


def test_persistent_claim_and_barriers_share_one_loop_level(device_asm):
    lines = device_asm
    starts = [i for i, l in enumerate(lines) if re.match(r'^_ZN12_GLOBAL__N_113k_tiles_queue.*:', l)]
    assert len(starts) >= 10       # 4 dtypes x matvec/grad + lookup + count
    for st in starts:
        end = next(i for i in range(st, len(lines)) if 's_endpgm' in lines[i])
        body = lines[st:end]
        claim = next(i for i, l in enumerate(body) if 'global_atomic_add v' in l and 'sc0' in l)
        bars = [i for i in range(claim, len(body)) if 's_barrier' in body[i]][:2]
        assert len(bars) == 2, lines[st]
        headers = [i for i in range(claim, bars[1]) if 'Loop Header' in body[i]]
        assert not headers, f'{lines[st]}: loop header between the claim and its barriers'
