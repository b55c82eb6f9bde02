"""Algorithmic work per unit of the trace kernels.

Counts are the arithmetic of the reference algorithm as written
(shared_shader/kerr_hamilton.h:97-178 for one damped Kerr RK4 step,
shared_shader/geodesic_math.h:101-134 for one Schwarzschild RK4 step), taken
by instantiating the CPU restatement with an operation-counting scalar
(tests/test_oracle_math.py::test_op_counts_are_stable keeps them in sync).
Flops are adds and multiplies (an FMA would count 2); divides and sin/cos are
counted separately and are not in the flop figure.
"""

KERR_STEP_OPS = {"add": 142, "mul": 348, "div": 53, "sqrt": 0, "trig": 9, "other": 0}
SCHW_STEP_OPS = {"add": 80, "mul": 162, "div": 29, "sqrt": 0, "trig": 8, "other": 0}

KERR_STEP_FLOPS = KERR_STEP_OPS["add"] + KERR_STEP_OPS["mul"]  # 490
KERR_STEP_DIVS = KERR_STEP_OPS["div"]  # 53
KERR_STEP_TRIG = KERR_STEP_OPS["trig"]  # 9
SCHW_STEP_FLOPS = SCHW_STEP_OPS["add"] + SCHW_STEP_OPS["mul"]  # 242

# MI355X published peaks
PEAK_FP32_TFLOPS = 157.3       # vector/matrix FP32 (packed issue)
PEAK_FP32_NONPACKED_TFLOPS = 78.6
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# Post passes: algorithmic HBM bytes per full-resolution pixel (RGBA16F HDR in,
# half-res RGBA16F bloom ping-pong, RGBA8 out):
#   extract 8 read + 2 write, blur_h 2 + 2, blur_v 2 + 2, composite 8 + 2 read + 4 write
POST_BYTES_PER_PIXEL = 32

# Measured HBM traffic of one mixed-precision FP32 trace dispatch at 1920x1080 x
# 16 spp (33 177 600 rays): rocprofv3 FETCH_SIZE 376424 KiB x 2.0 + WRITE_SIZE
# 522196 KiB x 1.0 (the factors are what the two counters read for a plain
# 256 MiB device-to-device copy), per traced ray: the set-up kernel's 16-byte
# init entry and 4-byte record index read, the 16-byte termination record,
# deferral markers and list entries written.
TRACE_HBM_BYTES_PER_RAY = (522196.0 * 1024 + 376424.0 * 1024 * 2.0) / 33177600.0

# VALU instructions the Kerr trace kernels issue per RK4 step and the lanes a
# wave64 instruction has active, from SQ counters of one 1080p x 16 spp mixed
# frame (one chunk): SQ_INSTS_VALU x (SQ_THREAD_CYCLES_VALU / SQ_ACTIVE_INST_VALU)
# / RK4 steps of the pass.  FP32 pass (trace_kernel<1, 1>, 64 steps between
# refills): 34.02 G instructions, 62.41 lanes, 7.704 G steps (34.13 G and 62.28
# lanes at 32 steps); FP64 list pass (trace_kernel<1, 3>, its list taken
# longest first, a refill every 16 steps): 1.320 G, 54.14 lanes, 0.2054 G
# steps (55.78 lanes and 1.285 G with a refill every 8; 47.65 lanes and
# 1.507 G instructions in record order).  A launch's issue rate is then steps x
# instructions per step / lanes / (launch time x 256 CUs x 2.4 GHz), against
# what the issue-rate calibration reaches: 1.9 wave64 FP32 instructions per CU
# per clock with all-vector operands (dual issue), 0.98 FP64 FMAs.
KERR_STEP_VALU_INSTS = 276
KERR_STEP_LANES = 62.41
KERR_STEP_VALU_INSTS_FP64 = 350
KERR_STEP_LANES_FP64 = 54.14
CU_COUNT = 256
SHADER_CLOCK_HZ = 2.4e9
VALU_ISSUE_PEAK_PER_CU_CLK = 1.9
VALU_ISSUE_PEAK_FP64_PER_CU_CLK = 0.98


def valu_issue_per_cu_clk(steps: float, launch_s: float, insts_per_step: float, lanes: float) -> float:
    """Wave64 VALU instructions issued per CU per clock by a launch of `steps` RK4 steps."""
    return steps * insts_per_step / lanes / (launch_s * CU_COUNT * SHADER_CLOCK_HZ)
