"""Register budget of the traversal kernels, read from the compiler's
resource-usage remarks for gfx950 (no GPU needed).

The closest-hit and visibility kernels are latency-bound: they are built for
8 waves per SIMD (64 VGPRs) with a 16-word LDS short stack, and their feature
specialisations must not pull in code the main BVH never runs (a call to an
out-of-line routine in one of them once took it to 194 VGPRs and 2 waves per
SIMD, doubling its run time).
"""
import os
import re
import shutil
import subprocess

import pytest

from luminara_amd import build as lbuild

KERNELS = os.path.join(os.path.dirname(lbuild.__file__), "csrc", "lmn_kernels.hip")


@pytest.fixture(scope="module")
def usage(tmp_path_factory):
    if not os.path.exists(lbuild.HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc is not available")
    out = tmp_path_factory.mktemp("ru") / "k.o"
    p = subprocess.run([lbuild.HIPCC, *lbuild.HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-c", KERNELS,
                        "-o", str(out)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert p.returncode == 0, p.stdout[-4000:]
    rows, cur = {}, None
    for line in p.stdout.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = m.group(1)
            rows[cur] = {}
            continue
        m = re.search(r"remark:\s+(VGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]): (\d+)", line)
        if m and cur:
            rows[cur][m.group(1)] = int(m.group(2))
    return rows


def traversal(usage):
    # mangled k_extend<SS, BS, FEAT> / k_shadow<SS, BS, FEAT>
    out = {}
    for name, r in usage.items():
        m = re.match(r"_ZN3lmn8(k_extend|k_shadow)ILi(\d+)ELi(\d+)ELj(\d+)E", name)
        if m:
            out[(m.group(1), int(m.group(2)), int(m.group(3)), int(m.group(4)))] = r
    return out


def test_traversal_kernels_keep_their_register_budget(usage):
    t = traversal(usage)
    assert len(t) >= 16, sorted(t)
    for (kern, ss, bs, feat), r in t.items():
        assert r["VGPRs"] <= 128, (kern, ss, bs, feat, r)
        if ss == 16:  # the 16-word LDS stack leaves room for 8 waves per SIMD
            assert r["VGPRs"] <= 64 and r["Occupancy [waves/SIMD]"] == 8, (kern, ss, bs, feat, r)


def test_batch_traversal_spills_little(usage):
    # the feature set of the triangle batch (no transforms, no FP64 records)
    t = traversal(usage)
    for kern in ("k_extend", "k_shadow"):
        r = t[(kern, 16, 256, 0)]
        assert r["ScratchSize [bytes/lane]"] <= 32, (kern, r)
