"""bench.py refuses more than one GPU with a message (one GPU is measured)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "one GPU" in out.stderr and "--gpus 1" in out.stderr


def test_reference_flop_model_of_the_config2_step():
    """performance_analyzer.f90:242-301 on the 16^3, 8-atom x 4-projector cell
    (123 sphere points per atom): 4*158*4096 + 4*(4 + 21*3936) + 4*8*4096."""
    sys.path.insert(0, ROOT)
    import bench
    from salmon_amd import synth
    s, _ = synth.config2(nsteps=2, orbitals=False)
    assert bench.ref_flop_per_orbital_step(s) == 2_588_672 + 330_640 + 131_072
