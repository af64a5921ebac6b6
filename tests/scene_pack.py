"""Builds a packed scene (the device layout the engine exports through
mrtxExportScene) on the CPU, for oracle path-tracer tests that run without a
GPU: one mesh, one or more instances, materials in GpuMaterial layout,
environment light first (reference Renderer.cpp:333-603 packing rules)."""
from __future__ import annotations

import numpy as np

from miniraytrace_amd.bvh import build_bvh


def _bounds(pos, idx):
    tri = pos[idx]  # (T, 3, 3)
    return np.concatenate([tri.min(1), tri.max(1)], axis=1).astype(np.float32)


def material(base=(0.8, 0.8, 0.8), opacity=1.0, emission=(0, 0, 0), roughness=0.5, metallic=0.0,
             transmission=0.0, ior=1.45):
    m = np.zeros(20, np.float32)
    m[0:3], m[3] = base, opacity
    m[4:7], m[7] = emission, roughness
    m[8], m[9], m[10] = metallic, transmission, ior
    m[12:17] = np.array([-1] * 5, np.int32).view(np.float32)
    return m


def pack(positions, indices, normals=None, materials=(), instances=((None, 0),), env=True):
    """instances: (xform3x4 or None, material index into `materials` + 1)."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1, 3)
    nodes, order = build_bvh(_bounds(pos, idx), 4)
    tris = np.zeros((idx.shape[0], 4), np.uint32)
    tris[:, :3] = idx[order]
    if normals is None:
        acc = np.zeros_like(pos, dtype=np.float64)
        fn = np.cross(pos[idx[:, 1]] - pos[idx[:, 0]], pos[idx[:, 2]] - pos[idx[:, 0]])
        for k in range(3):
            np.add.at(acc, idx[:, k], fn)
        normals = acc / np.maximum(np.linalg.norm(acc, axis=1, keepdims=True), 1e-30)
    P = np.zeros((pos.shape[0], 4), np.float32)
    P[:, :3] = pos
    N = np.zeros((pos.shape[0], 4), np.float32)
    N[:, :3] = normals
    mats = [material()] + list(materials)
    inst_rows, inst_bounds = [], []
    lo, hi = pos.min(0), pos.max(0)
    for xf, mat in instances:
        m = np.eye(4, dtype=np.float32)
        if xf is not None:
            m[:3, :] = np.asarray(xf, np.float32).reshape(3, 4)
        inv = np.linalg.inv(m.astype(np.float64)).astype(np.float32)
        row = np.zeros(36, np.float32)
        row[0:16] = m.T.reshape(-1)   # column-major
        row[16:32] = inv.T.reshape(-1)
        row[32] = np.uint32(0).view(np.float32)
        row[33] = np.uint32(mat).view(np.float32)
        inst_rows.append(row)
        corners = np.array([[x, y, z, 1] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])],
                           np.float32) @ m.T
        inst_bounds.append(np.concatenate([corners[:, :3].min(0), corners[:, :3].max(0)]))
    tl_nodes, tl_order = build_bvh(np.array(inst_bounds, np.float32), 1)
    tl_nodes = tl_nodes.copy()
    internal = (tl_nodes["packed"] >> 2) == 0
    tl_nodes["leftOrPrim"][internal] += len(nodes)
    inst = np.array(inst_rows, np.float32)[tl_order]
    lights = np.zeros((1 if env else 0, 20), np.float32)
    return {
        "nodes": np.concatenate([nodes, tl_nodes]),
        "tlasRoot": len(nodes),
        "triangles": tris,
        "positions": P,
        "normals": N,
        "instances": inst,
        "materials": np.array(mats, np.float32),
        "lights": lights,
        "envAlias": np.zeros((0, 2), np.uint32),
    }
