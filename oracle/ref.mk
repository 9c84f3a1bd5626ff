# Builds the unmodified Osprey-DPD reference (C++ sources under /root/reference/src)
# into oracle/_ref/dpd, compiling each source file with g++ directly as the
# reference README describes (g++ -c -O3 -std=c++11 *.cpp; g++ -o dpd *.o).
# The binary is used only as a checker: it writes golden fixtures in this
# container and is timed as the CPU baseline. Nothing in the product links it.
#
#   make -f oracle/ref.mk -j8          (from the repository root)

REF_SRC ?= /root/reference/src
OUT     := oracle/_ref
OBJDIR  := $(OUT)/obj
SRCS    := $(wildcard $(REF_SRC)/*.cpp)
OBJS    := $(patsubst $(REF_SRC)/%.cpp,$(OBJDIR)/%.o,$(SRCS))
CXX     ?= g++
CXXFLAGS := -O3 -std=c++11 -w

all: $(OUT)/dpd

$(OUT)/dpd: $(OBJS)
	$(CXX) -o $@ $(OBJS) -pthread

$(OBJDIR)/%.o: $(REF_SRC)/%.cpp | $(OBJDIR)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OBJDIR):
	mkdir -p $@

.PHONY: all
