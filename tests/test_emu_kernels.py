"""The library's kernels run on the CPU through the launch emulation in
tests/emu (work-groups of fibers, host buffers as device memory), checked
against the oracle at small sizes: single and batched NTT with plans and
graph replay, the gate, single and batched MSM over shared table sets.  This
covers the kernels' indexing and pass logic without a GPU; the arithmetic is
the compiler-scheduled multiply the library is built with."""
import shutil

import pytest

from tests.emu import build as B
from tests.emu import scenarios as S

pytestmark = pytest.mark.skipif(not B.CLANG or not shutil.which(B.CLANG) and
                                not __import__("os").path.exists(B.CLANG),
                                reason="no host clang++ to build the emulation")


@pytest.fixture(scope="module")
def emu():
    B.build()
    from tests.emu.harness import Emu
    return Emu()


@pytest.mark.parametrize("name", [n for n, _ in S.ALL])
def test_scenario(emu, name):
    dict(S.ALL)[name](emu)
