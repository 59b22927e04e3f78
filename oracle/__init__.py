"""CPU oracle — TEST INFRASTRUCTURE ONLY.

A restatement of the reference algorithms (CosmicBLUP, C++/Eigen) for the
GBLUP/ssGBLUP hot path, used to check the HIP path and to time a CPU
baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg may import it. The product library never calls it.
"""
