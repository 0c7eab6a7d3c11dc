"""ctypes declarations of include/mi355sssp.h and include/mi355graph.h."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.normpath(os.path.join(_HERE, "..", "lib"))

ALG_BF, ALG_BF_JACOBI, ALG_DELTA, ALG_ADDS = 0, 1, 2, 3
INF_IEEE, INF_FLT_MAX = 0, 1
W_F32, W_U32 = 0, 1
ORDER_AUTO, ORDER_INPUT, ORDER_DEGREE = 0, 1, 2
OUT, IN = 0, 1
HOST, DEVICE = 0, 1

STATUS_NAMES = {
    0: "MI355_OK", -1: "MI355_EINVAL", -2: "MI355_ERANGE", -3: "MI355_EOVERFLOW",
    -4: "MI355_EHIP", -5: "MI355_ENOMEM", -6: "MI355_ECAPACITY", -7: "MI355_EDEVICE",
}


class Mi355Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class CsrDesc(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("nnz", C.c_int64),
                ("rowptr", C.c_void_p), ("col", C.c_void_p), ("w", C.c_void_p),
                ("offset_bits", C.c_int32), ("weight_type", C.c_int32),
                ("orientation", C.c_int32), ("memory", C.c_int32)]


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("weight_type", C.c_int32),
                ("max_degree", C.c_int32), ("avg_degree", C.c_double),
                ("avg_weight", C.c_double), ("large_vertices", C.c_int64),
                ("large_chunks", C.c_int64), ("has_negative", C.c_int32), ("device", C.c_int32),
                ("vertex_order", C.c_int32), ("reserved", C.c_int32)]


class SsspOptions(C.Structure):
    _fields_ = [("algorithm", C.c_int32), ("max_iters", C.c_int32),
                ("check_interval", C.c_int32), ("delta", C.c_float), ("inf_mode", C.c_int32),
                ("dynamic_delta", C.c_int32), ("use_mlp", C.c_int32),
                ("worker_blocks", C.c_int32), ("rounds_per_sync", C.c_int32),
                ("active_buckets", C.c_int32), ("pull", C.c_int32), ("pull_margin", C.c_float),
                ("light_scale", C.c_float), ("reserved", C.c_int32 * 3)]


class SsspStats(C.Structure):
    _fields_ = [("iterations_used", C.c_int32), ("converged", C.c_int32), ("rounds", C.c_int32),
                ("syncs", C.c_int32), ("edges_relaxed", C.c_int64),
                ("vertices_processed", C.c_int64), ("reached", C.c_int64),
                ("reached_edges", C.c_int64), ("initial_delta", C.c_float),
                ("final_delta", C.c_float), ("device_ms", C.c_float), ("kernel_ms", C.c_float),
                ("has_negative_cycle", C.c_int32), ("policy", C.c_int32),
                ("pulled_phases", C.c_int32), ("reserved", C.c_int32)]


class MlpParams(C.Structure):
    _fields_ = [("w1", C.c_float * (64 * 32)), ("b1", C.c_float * 64), ("w2", C.c_float * (4 * 64)),
                ("b2", C.c_float * 4), ("feat_mean", C.c_float * 32), ("feat_inv_std", C.c_float * 32),
                ("graph_feat_lo", C.c_float * 14), ("graph_feat_hi", C.c_float * 14)]


class MlpTrainOptions(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("weight_decay", C.c_float),
                ("seed", C.c_uint64)]


ADDS_TRACE_STRIDE = 40


class QueryBatch(C.Structure):
    _fields_ = [("grid_cols", C.c_int32), ("count", C.c_int32), ("src_offsets", C.c_void_p),
                ("sources", C.c_void_p), ("boxes", C.c_void_p)]


class BatchOptions(C.Structure):
    _fields_ = [("rounds_per_graph", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("use_graph", C.c_int32), ("inf_mode", C.c_int32), ("count_edges", C.c_int32),
                ("order", C.c_int32), ("delta_scale", C.c_float), ("reserved", C.c_int32),
                ("queue_capacity", C.c_int64)]


class BatchStats(C.Structure):
    _fields_ = [("label_slots", C.c_int64), ("sources", C.c_int64), ("rounds", C.c_int32),
                ("graph_replays", C.c_int32), ("rounds_per_graph", C.c_int32),
                ("far_passes", C.c_int32), ("edges_relaxed", C.c_int64),
                ("queue_capacity", C.c_int64), ("device_ms", C.c_float), ("stopped_queries", C.c_int32)]


class BfProgress(C.Structure):
    _fields_ = [("iteration", C.c_int), ("max_iters", C.c_int),
                ("convergence_checked", C.c_ubyte), ("changed", C.c_ubyte)]


PROGRESS_CB = C.CFUNCTYPE(None, C.POINTER(BfProgress), C.c_void_p)


class BfCsrStats(C.Structure):
    _fields_ = [("iterations_used", C.c_int), ("converged", C.c_int),
                ("check_interval", C.c_int), ("checks_performed", C.c_int),
                ("changed_checks", C.c_int), ("unchanged_checks", C.c_int)] + [
        (name, C.c_double) for name in (
            "initialization_ms", "bellman_ford_loop_ms", "relaxation_ms", "checked_merge_ms",
            "unchecked_merge_ms", "checked_iteration_ms", "unchecked_iteration_ms",
            "final_copy_ms", "total_runtime_ms", "estimated_check_overhead_ms")]


class GView(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("nnz", C.c_int64),
                ("weight_type", C.c_int32), ("orientation", C.c_int32),
                ("rowptr", C.POINTER(C.c_int64)), ("col", C.POINTER(C.c_int32)),
                ("w", C.POINTER(C.c_uint32)), ("declared_edges", C.c_uint64),
                ("loaded_edges", C.c_uint64)]


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() first")
    return C.CDLL(path)


_engine = None
_graph = None

P = C.c_void_p
I32, I64, U32, U64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64

_ENGINE_SIGS = {
    "mi355_last_error": (C.c_char_p, []),
    "mi355_last_hip_error": (C.c_int, []),
    "mi355_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mi355_set_device": (C.c_int, [C.c_int]),
    "mi355_device_malloc": (C.c_int, [C.POINTER(P), C.c_size_t]),
    "mi355_device_free": (C.c_int, [P]),
    "mi355_memcpy_h2d": (C.c_int, [P, P, C.c_size_t, P]),
    "mi355_memcpy_d2h": (C.c_int, [P, P, C.c_size_t, P]),
    "mi355_memset_d8": (C.c_int, [P, C.c_int, C.c_size_t, P]),
    "mi355_device_gather_u64": (C.c_int, [P, P, P, C.c_int64, P]),
    "mi355_memcpy_d2d": (C.c_int, [P, P, C.c_size_t, P]),
    "mi355_stream_create": (C.c_int, [C.POINTER(P)]),
    "mi355_stream_destroy": (C.c_int, [P]),
    "mi355_stream_sync": (C.c_int, [P]),
    "mi355_graph_create": (C.c_int, [C.POINTER(CsrDesc), I32, P, C.POINTER(P)]),
    "mi355_graph_create_ex": (C.c_int, [C.POINTER(CsrDesc), I32, I32, P, C.POINTER(P)]),
    "mi355_graph_destroy": (C.c_int, [P]),
    "mi355_graph_info_get": (C.c_int, [P, C.POINTER(GraphInfo)]),
    "mi355_graph_download": (C.c_int, [P, P, P, P]),
    "mi355_sssp_options_default": (None, [C.POINTER(SsspOptions)]),
    "mi355_sssp": (C.c_int, [P, I32, C.POINTER(SsspOptions), P, P, C.POINTER(SsspStats), P]),
    "mi355_sssp_labels": (C.c_int, [P, I32, C.POINTER(SsspOptions), C.POINTER(SsspStats), P]),
    "mi355_sssp_labels_to_output": (C.c_int, [P, I32, I32, P, P, P]),
    "mi355_sssp_round_log": (C.c_int, [P, P, I32]),
    "mi355_bf_csr_f32": (C.c_int, [C.POINTER(CsrDesc), C.c_int, C.c_int, C.c_int, P,
                                   C.POINTER(BfCsrStats), PROGRESS_CB, P, P]),
    "mi355_delta_stepping_csr_f32": (C.c_int, [C.POINTER(CsrDesc), C.c_int, C.c_float, C.c_int, P,
                                               C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               PROGRESS_CB, P, P]),
    "mi355_batch_options_default": (None, [C.POINTER(BatchOptions)]),
    "mi355_batch_layout": (C.c_int, [P, C.POINTER(QueryBatch), P]),
    "mi355_sssp_batched": (C.c_int, [P, C.POINTER(QueryBatch), C.POINTER(BatchOptions), P, P,
                                     C.POINTER(BatchStats), P]),
    "mi355_sssp_batched_targets": (C.c_int, [P, C.POINTER(QueryBatch), P, P, C.POINTER(BatchOptions), P, P,
                                             C.POINTER(BatchStats), P]),
    "mi355_batch_paths": (C.c_int, [P, C.POINTER(QueryBatch), P, P, P, P, P, I64, P, P, P, I64, P]),
    "mi355_batch_autotune": (C.c_int, [P, C.POINTER(QueryBatch), C.c_int32, C.POINTER(BatchOptions),
                                       P, P]),
    "mi355_mlp_train_options_default": (None, [C.POINTER(MlpTrainOptions)]),
    "mi355_mlp_forward": (C.c_int, [C.POINTER(MlpParams), P, I64, P]),
    "mi355_mlp_loss_grad": (C.c_int, [C.POINTER(MlpParams), P, P, P, I64, C.POINTER(C.c_float),
                                      C.POINTER(MlpParams)]),
    "mi355_mlp_train": (C.c_int, [C.POINTER(MlpParams), P, P, P, I64, C.POINTER(MlpTrainOptions), P]),
    "mi355_adds_mlp_set": (C.c_int, [C.POINTER(MlpParams)]),
    "mi355_adds_default_workers": (C.c_int, [C.POINTER(C.c_int32)]),
    "mi355_adds_mlp_clear": (C.c_int, []),
    "mi355_adds_schedule_set": (C.c_int, [P, I32]),
    "mi355_adds_trace": (C.c_int, [P, I32, C.POINTER(SsspOptions), P, I32, C.POINTER(I32),
                                   C.POINTER(SsspStats), P]),
}

_GRAPH_SIGS = {
    "mi355g_last_error": (C.c_char_p, []),
    "mi355g_view": (C.c_int, [P, C.POINTER(GView)]),
    "mi355g_free": (None, [P]),
    "mi355g_from_csr": (C.c_int, [I64, I64, I64, P, P, P, I32, I32, C.POINTER(P)]),
    "mi355g_from_edges": (C.c_int, [I64, I64, P, P, P, I32, I32, I32, C.POINTER(P)]),
    "mi355g_transpose": (C.c_int, [P, C.POINTER(P)]),
    "mi355g_read_ripscsr1": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "mi355g_write_ripscsr1": (C.c_int, [P, C.c_char_p]),
    "mi355g_write_ripsadj1": (C.c_int, [P, C.c_char_p]),
    "mi355g_read_ripsadj1": (C.c_int, [C.c_char_p, P, I64, C.POINTER(I64)]),
    "mi355g_from_dense": (C.c_int, [P, I64, I32, I32, I32, C.POINTER(P)]),
    "mi355g_read_dimacs": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "mi355g_read_galois_gr": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "mi355g_write_galois_gr": (C.c_int, [P, C.c_char_p]),
    "mi355g_gen_grid4": (C.c_int, [I32, I32, U32, U32, U64, C.POINTER(P)]),
    "mi355g_gen_rmat": (C.c_int, [I32, I32, C.c_double, C.c_double, C.c_double, U64, I32,
                                  C.POINTER(P)]),
    "mi355g_gen_routing_grid": (C.c_int, [I32, C.c_double, C.c_double, C.c_double, C.c_double,
                                          U64, C.POINTER(P)]),
    "mi355g_gen_chain_random": (C.c_int, [I64, I32, C.c_double, C.c_double, I32, U64,
                                          C.POINTER(P)]),
    "mi355g_pick_sources": (C.c_int, [P, I32, U64, P]),
    "mi355g_sha256": (C.c_int, [P, C.c_char_p]),
    "mi355g_rng_u64": (U64, [U64, U64, U64]),
    "mi355g_rng_unit_f32": (C.c_float, [U64, U64, U64]),
    "mi355g_fill_unit_f32": (None, [U64, U64, U64, I64, P]),
    "mi355g_read_ripsifm1": (C.c_int, [C.c_char_p, C.POINTER(P)]),
    "mi355g_ifm_free": (None, [P]),
    "mi355g_ifm_header": (C.c_int, [P, P]),
    "mi355g_ifm_string": (C.c_int, [P, U64, C.POINTER(C.c_void_p), C.POINTER(U64)]),
    "mi355g_ifm_tables": (C.c_int, [P, P, P, P, P, P, P]),
    "mi355g_ifm_requests": (C.c_int, [P, P, P, P, P, P]),
    "mi355g_ifm_logical": (C.c_int, [P, P, P, P]),
    "mi355g_ifm_payloads": (C.c_int, [P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
}

IFM_HEADER_FIELDS = ("version", "orientation", "string_count", "node_count", "edge_attr_count",
                     "pip_data_count", "site_pin_attr_count", "route_request_count",
                     "blocked_node_count", "sink_stop_node_count", "logical_cell_count",
                     "logical_net_count", "logical_port_instance_count",
                     "physical_netlist_byte_count", "logical_netlist_byte_count",
                     "device_path_string", "physical_path_string", "logical_path_string",
                     "logical_design_name_string", "route_source_count", "route_sink_count")


class IfmHeader(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in IFM_HEADER_FIELDS]


def _bind(lib: C.CDLL, sigs: dict) -> C.CDLL:
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def engine() -> C.CDLL:
    global _engine
    if _engine is None:
        _engine = _bind(_load("libmi355sssp.so"), _ENGINE_SIGS)
    return _engine


def graphlib() -> C.CDLL:
    global _graph
    if _graph is None:
        _graph = _bind(_load("libmi355graph.so"), _GRAPH_SIGS)
    return _graph


def check(status: int) -> None:
    if status != 0:
        raise Mi355Error(status, engine().mi355_last_error().decode())


def gcheck(status: int) -> None:
    if status != 0:
        raise Mi355Error(status, graphlib().mi355g_last_error().decode())


class OneOptions(C.Structure):
    _fields_ = [("max_iters", C.c_int), ("num_buckets", C.c_int),
                ("bucket_capacity_per_bucket", C.c_int), ("initial_delta", C.c_float),
                ("dynamic_delta", C.c_int), ("target_low_work", C.c_int),
                ("target_high_work", C.c_int)]


class OneResult(C.Structure):
    _fields_ = [("iterations_used", C.c_int), ("converged", C.c_int),
                ("has_negative_cycle", C.c_int), ("work_count", C.c_ulonglong),
                ("milliseconds", C.c_float), ("final_delta", C.c_float)]


class AddsOptions(C.Structure):
    _fields_ = [("worker_blocks", C.c_uint32), ("assignment_chunk", C.c_uint32),
                ("initial_active_buckets", C.c_uint32), ("max_active_buckets", C.c_uint32),
                ("initial_delta", C.c_double), ("initial_delta_scale", C.c_double),
                ("min_delta", C.c_double), ("delta_growth", C.c_double),
                ("delta_shrink", C.c_double), ("clip_threshold", C.c_double),
                ("utilization_low", C.c_double), ("utilization_high", C.c_double),
                ("adjust_period_head_switches", C.c_uint32), ("dynamic_delta", C.c_int32),
                ("dynamic_active_buckets", C.c_int32), ("initialize_distances", C.c_int32),
                ("pool_words", C.c_uint64), ("memory_fraction", C.c_double)]

    @classmethod
    def defaults(cls):
        o = cls()
        o.assignment_chunk, o.initial_active_buckets, o.max_active_buckets = 256, 1, 32
        o.initial_delta_scale, o.min_delta = 16.0, 1e-3
        o.delta_growth, o.delta_shrink, o.clip_threshold = 1.5, 0.75, 0.65
        o.utilization_low, o.utilization_high = 0.35, 0.85
        o.adjust_period_head_switches = 8
        o.dynamic_delta = o.dynamic_active_buckets = o.initialize_distances = 1
        o.memory_fraction = 0.45
        return o


class AddsStats(C.Structure):
    _fields_ = [("initial_delta", C.c_double), ("final_delta", C.c_double),
                ("final_bucket_base", C.c_double), ("worker_blocks", C.c_uint32),
                ("final_active_buckets", C.c_uint32), ("pool_pages", C.c_uint32),
                ("device_error", C.c_uint32)] + [
        (n, C.c_uint64) for n in ("assigned_items", "processed_items", "enqueued_items",
                                  "edge_relaxations", "successful_relaxations", "head_switches",
                                  "delta_updates", "max_inflight_items")]


_ENGINE_SIGS.update({
    "mi355_minplus_gemm_f32": (C.c_int, [P, P, P] + [C.c_int] * 6 + [P]),
    "mi355_minplus_gemm_f32_batched": (C.c_int, [C.c_int] + [P] * 9 + [P]),
    "mi355_minplus_gemm_f32_batched_noalloc": (C.c_int, [C.c_int] + [P] * 9 + [P]),
    "mi355_minplus_tile_count": (C.c_int, []),
    "mi355_minplus_tile_name": (C.c_char_p, [C.c_int]),
    "mi355_minplus_autotune": (C.c_int, [P, C.c_int, P, P]),
    "mi355_minplus_tuning_get": (C.c_int, [P]),
    "mi355_minplus_tuning_set": (C.c_int, [P]),
    "mi355_minplus_spgemm_csr_f32": (C.c_int, [I64, I64, I64, P, P, P, I64, I64, I64, P, P, P,
                                               C.c_int, C.c_int, P, P, P, P, P]),
    "mi355_sparse_vector_min_merge": (C.c_int, [I64, P, P, P, P, P, P, P, P, P, P]),
    "mi355_bf_dense_f32": (C.c_int, [P, I32, C.c_int, C.c_int, C.c_int, P, P, P, P]),
    "mi355_bf_gauss_seidel_f32": (C.c_int, [P, I32, C.c_int, C.c_int, C.c_int, P, P, P]),
    "mi355_one_sssp_f32": (C.c_int, [I32, C.c_int, C.c_int, P, P, P, C.c_int,
                                     C.POINTER(OneOptions), P, C.POINTER(OneResult), P]),
    "mi355_hip_adds_sssp": (C.c_int, [I32, U32, U32, P, P, P, U32, P, C.POINTER(AddsStats),
                                      C.POINTER(AddsOptions), P]),
})
