"""Copy the reference's committed golden outputs (data files only) into
tests/golden/.  Run once in a container that has the reference checkout at
/root/reference; the tests read only the copies.

Sources (reference repository paths):
  scripts/negf_python/PointContact/cpp_results/1gpu/{Sigma,DOS,Transmission}
      spectral_function_1D_pll harness: N=4 chain, (17,0) mode J=6, 1000
      energies on [-1,1] eV with Im E = 1e-8, U = 0.
  scripts/negf_python/DOS_Compare/cpp_results/gpu/{LDOS.dat,Transmission.dat}
      main-code DOS path, all (17,0) modes, one unit cell (N=4).
  scripts/negf_cpp/decimation_technique/Exec/output/{Sigma,DOS,Transmission}
      decimation harness: N=4, 10 energies on [-6,0] eV, Im E = 1e-8.
"""
import os
import shutil

SRC = "/root/reference/scripts"
DST = os.path.dirname(os.path.abspath(__file__))
FILES = {
    "pointcontact_Sigma.txt": "negf_python/PointContact/cpp_results/1gpu/Sigma",
    "pointcontact_DOS.txt": "negf_python/PointContact/cpp_results/1gpu/DOS",
    "pointcontact_Transmission.txt": "negf_python/PointContact/cpp_results/1gpu/Transmission",
    "doscompare_LDOS.txt": "negf_python/DOS_Compare/cpp_results/gpu/LDOS.dat",
    "doscompare_Transmission.txt": "negf_python/DOS_Compare/cpp_results/gpu/Transmission.dat",
    "decimation_Sigma.txt": "negf_cpp/decimation_technique/Exec/output/Sigma",
    "decimation_DOS.txt": "negf_cpp/decimation_technique/Exec/output/DOS",
    "decimation_Transmission.txt": "negf_cpp/decimation_technique/Exec/output/Transmission",
}

if __name__ == "__main__":
    for dst, src in FILES.items():
        shutil.copyfile(os.path.join(SRC, src), os.path.join(DST, dst))
        print(dst, "<-", src)
