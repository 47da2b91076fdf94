"""windtunnel.profsum: the HBM-traffic summary scales FETCH_SIZE / WRITE_SIZE by
a copy of known size and prices the fused step + advection launch at its cells
and particles (no GPU: synthetic rocprofv3 counter CSVs)."""
import csv
import json
import os

FIELDS = ['Kernel_Name', 'Grid_Size', 'Counter_Name', 'Counter_Value']


def _write(d, counter, rows):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, 'r_counter_collection.csv'), 'w', newline='') as fh:
        w = csv.DictWriter(fh, fieldnames=FIELDS)
        w.writeheader()
        for name, grid, v in rows:
            w.writerow({'Kernel_Name': name, 'Grid_Size': grid, 'Counter_Name': counter, 'Counter_Value': v})


def test_traffic_scaling_and_fused_pricing(tmp_path):
    from windtunnel import profsum
    cells, parts, copy = 1000, 50, 10 ** 6
    # the copy reads and writes 1e6 B, counted as 500 / 1000 units (medians) -> 2000 / 1000 B per unit
    fetch = [('void wt::k_stream_copy(...)', 4096, 500.0), ('void wt::k_stream_copy(...)', 4096, 510.0),
             ('void wt::k_stream_copy(...)', 4096, 490.0),
             ('void wt::k_step<0, 1>(wt::StepDesc, int, int, int)', 1024, 40.0),
             ('void wt::k_step_adv<0, 1>(wt::StepDesc, wt::AdvectArgs, int, int, int, unsigned int)', 1280, 42.0),
             ('void wt::k_step_batch<1, 1>(wt::StepDesc const*, int, int, int, int)', 2048, 7.0)]
    write = [(n, g, v * 2) for n, g, v in fetch]
    _write(tmp_path / 'f', 'FETCH_SIZE', fetch)
    _write(tmp_path / 'w', 'WRITE_SIZE', write)
    out = profsum.traffic(str(tmp_path / 'f'), str(tmp_path / 'w'), copy, str(tmp_path), cells, parts)
    # batched steps get a table row priced over all cells (FP16: 85 B) but no json
    row = [ln for ln in out.splitlines() if 'k_step_batch' in ln]
    assert len(row) == 1 and row[0].endswith(f'| {28000 / (85 * cells):.4f} |')
    assert not os.path.exists(tmp_path / 'step_fp16_512x256x256_traffic.json')
    plain = json.load(open(tmp_path / 'step_fp32_512x256x256_traffic.json'))
    fused = json.load(open(tmp_path / 'step_adv_fp32_512x256x256_traffic.json'))
    assert plain['fetch_bytes'] == 40.0 * 2000 and plain['write_bytes'] == 80.0 * 1000
    assert plain['algorithmic_bytes_per_launch'] == 169 * cells and plain['particles'] == 0
    assert fused['fetch_bytes'] == 42.0 * 2000 and fused['write_bytes'] == 84.0 * 1000
    assert fused['algorithmic_bytes_per_launch'] == 169 * cells + 40 * parts
    assert fused['hbm_bytes_per_launch'] == fused['fetch_bytes'] + fused['write_bytes']


def test_traffic_batch_with_copy_from_other_runs_and_no_mac(tmp_path):
    """A batch's counter runs carry no copy: the scale comes from other runs on the
    box (--copy-dirs); --no-mac prices the steps without the macroscopic store."""
    from windtunnel import profsum
    cells, copy = 1000, 10 ** 6
    _write(tmp_path / 'cf', 'FETCH_SIZE', [('void wt::k_stream_copy(...)', 4096, 500.0)])
    _write(tmp_path / 'cw', 'WRITE_SIZE', [('void wt::k_stream_copy(...)', 4096, 1000.0)])
    batch = [('void wt::k_step_batch<0, 1>(wt::StepDesc const*, int, int, int, int)', 2048, 60.0)]
    _write(tmp_path / 'f', 'FETCH_SIZE', batch)
    _write(tmp_path / 'w', 'WRITE_SIZE', [(n, g, 2 * v) for n, g, v in batch])
    out = profsum.traffic(str(tmp_path / 'f'), str(tmp_path / 'w'), copy, None, cells,
                          copy_dirs=(str(tmp_path / 'cf'), str(tmp_path / 'cw')), no_mac=True)
    # 60 x 2000 + 120 x 1000 = 240000 B against (169 - 16) x 1000
    row = [ln for ln in out.splitlines() if 'k_step_batch' in ln][0]
    assert row.endswith(f'| {240000 / (153 * cells):.4f} |')
