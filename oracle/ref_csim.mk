# Compiles the reference's own FP64 csim kernels (src/csim) into
# $(OUT)/libcsim_ref.so, straight from the reference tree, to pin the oracle's
# gate kernels.  No reference source is copied into this repository.
# Usage: make -f oracle/ref_csim.mk REF=/root/reference OUT=oracle/_ref
REF ?= /root/reference
OUT ?= oracle/_ref
CSIM := $(REF)/src/csim
SRCS := constant.cpp utility.cpp init_ops_fill.cpp stat_ops.cpp \
        update_ops_matrix_dense_single.cpp update_ops_matrix_dense_double.cpp \
        update_ops_matrix_dense_multi.cpp update_ops_named_CNOT.cpp update_ops_named_CZ.cpp \
        update_ops_named_H.cpp update_ops_named_X.cpp update_ops_named_SWAP.cpp \
        update_ops_pauli_single.cpp update_ops_named_state.cpp \
        update_ops_control_single_target_single.cpp update_ops_control_multi_target_single.cpp \
        update_ops_control_multi_target_multi.cpp update_ops_control_single_target_multi.cpp \
        update_ops_matrix_diagonal_single.cpp update_ops_named_Y.cpp update_ops_named_Z.cpp
OBJS := $(addprefix $(OUT)/,$(SRCS:.cpp=.o))
CXXFLAGS := -O2 -fPIC -fopenmp -std=c++14 -I$(REF)/src -w

$(OUT)/libcsim_ref.so: $(OBJS)
	g++ -shared -fopenmp -o $@ $(OBJS)

$(OUT)/%.o: $(CSIM)/%.cpp
	@mkdir -p $(OUT)
	g++ $(CXXFLAGS) -c $< -o $@
