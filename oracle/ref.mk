# Builds the reference's own FP64 CPU NEP evaluator into oracle/_ref (test
# infrastructure only).  The reference source is compiled where it lies; no
# reference file is copied into this repository.  nep_cpu.cpp needs <mpi.h>
# and MPI_Allreduce (charge path only): both come from the MPICH that ships in
# /opt/conda.  Usage: make -f oracle/ref.mk   (from the repository root)
REF ?= /root/reference/third_party/lammps-matpl-overlay/files/src
MPI_INC ?= /opt/conda/include
MPI_LIB ?= /opt/conda/lib
CXX ?= g++
OUT := oracle/_ref/libnep_ref.so

all: $(OUT)

$(OUT): $(REF)/nep_cpu.cpp $(REF)/nep_cpu.h oracle/ref_driver.cpp
	mkdir -p oracle/_ref
	$(CXX) -O3 -march=x86-64-v3 -std=c++17 -fPIC -shared -w -I$(REF) -I$(MPI_INC) \
	  $(REF)/nep_cpu.cpp oracle/ref_driver.cpp -L$(MPI_LIB) -lmpi -Wl,-rpath,$(MPI_LIB) -o $(OUT)
