"""Readback corruptions for the host trust boundary, one per invariant class of
the reference's fault decorator (src/gpu/fault_injecting_backend.cc:124-224),
applied to the first query of a compact-path readback.  The workspace faults
(kWorkspaceBounds, kWorkspaceOwner) act on the full-record readback and have
no compact-form counterpart."""
import ctypes as C

from apgar_amd import v1

INVALID = 0xFFFFFFFF


def _copy(struct):
    out = type(struct)()
    C.memmove(C.addressof(out), C.addressof(struct), C.sizeof(struct))
    return out


def _path_slice(paths, states):
    p = paths[0]
    return p.state_offset, p.state_count


def inject(fault, results, paths, states, telemetry):
    """Returns corrupted copies (results, paths, states, telemetry) and the
    invariant id the host validator must report."""
    results = [_copy(r) for r in results]
    paths = [_copy(p) for p in paths]
    states = list(states)
    telemetry = _copy(telemetry)
    r = results[0]
    off, n = _path_slice(paths, states)
    if fault == "query_telemetry":
        r.rounds = telemetry.dispatched_rounds + 1
        return results, paths, states, telemetry, "gpu.batch.query.telemetry.v1"
    if fault == "memory_accounting":
        telemetry.batch_host_bytes += 1
        return results, paths, states, telemetry, "gpu.batch.memory.accounting.v1"
    if fault == "batch_telemetry":
        telemetry.kernel_launch_count += 1
        return results, paths, states, telemetry, "gpu.batch.telemetry.v1"
    if fault == "query_identity":
        r.query_id += 1
        return results, paths, states, telemetry, "gpu.batch.query.identity.v1"
    if fault == "false_disconnected":
        r.completion = v1.DISCONNECTED
        r.goal_state = INVALID
        # drop the path's states and fix the byte accounting, as a real
        # producer of the claim would
        del states[off:off + n]
        for p in paths[1:]:
            if p.state_offset > off:
                p.state_offset -= n
        paths[0].state_count = 0
        paths[0].state_offset = 0
        telemetry.batch_host_bytes -= 4 * n
        telemetry.device_to_host_readback_bytes -= 4 * n
        return results, paths, states, telemetry, "gpu.disconnected.cpu_reachable.v1"
    if fault == "compact_path_bounds":
        paths[0].schema_version += 1
        return results, paths, states, telemetry, "gpu.batch.compact_path.bounds.v1"
    if fault == "compact_path_endpoint":
        states[off] = INVALID
        return results, paths, states, telemetry, "gpu.batch.compact_path.endpoint.v1"
    if fault == "compact_path_cycle":
        assert n >= 4
        states[off + 1] = states[off + 2]
        return results, paths, states, telemetry, "gpu.batch.compact_path.cycle.v1"
    if fault == "compact_path_heading":
        start = r.start_node * 9 + 8
        cur = states[off:off + n]
        for i in range(1, n - 1):
            node = cur[i] // 9
            rep = node * 9 + 8
            if node != r.start_node and rep not in cur and rep != start:
                states[off + i] = rep
                break
        else:
            raise AssertionError("no interior state to corrupt")
        return results, paths, states, telemetry, "gpu.batch.compact_path.heading.v1"
    if fault == "compact_path_edge":
        cur = states[off:off + n]
        for i in range(1, n - 1):
            node, h = divmod(cur[i], 9)
            for nh in range(8):
                rep = node * 9 + nh
                if nh != h and rep not in cur:
                    states[off + i] = rep
                    return results, paths, states, telemetry, "gpu.batch.compact_path.edge.v1"
        raise AssertionError("no interior state to corrupt")
    raise KeyError(fault)


FAULTS = ["query_telemetry", "memory_accounting", "batch_telemetry", "query_identity", "false_disconnected",
          "compact_path_bounds", "compact_path_endpoint", "compact_path_cycle", "compact_path_heading",
          "compact_path_edge"]
