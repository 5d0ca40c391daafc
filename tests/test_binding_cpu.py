"""The cpGPUSpace / cpGPUBatch binding's host logic over the oracle's CPU step
(oracle/_ref/libcpgpuspace_cpu.so: the binding compiled with
oracle/cpgpu_on_oracle.c in place of the device library) -- no GPU needed.

The same flows as tests/test_gpuspace.py (tests/binding_flows.py): packing
Chipmunk's structs, uploading setters, packing again after adds and removes
with the surviving warm starts, threading the arbiters onto the bodies, and
a batch of spaces; each against an independent oracle stepping the binding's
own description.
"""
import pytest

import binding_flows as F
import gpuspace

pytestmark = pytest.mark.skipif(not gpuspace.available("cpu"),
                                reason="the binding is built only where the reference sources are present")


@pytest.fixture(autouse=True)
def _errors_raise():
    F.errors_raise("cpu")
    yield
    F.errors_abort("cpu")


def test_steps_and_threaded_arbiters():
    F.steps_bit_exact_and_threads_arbiters("cpu")


def test_setters_between_steps():
    F.setters_between_steps_reach_the_world("cpu")


def test_add_and_remove_keep_warm_starts():
    F.add_and_remove_keep_warm_starts("cpu")


def test_batch_of_plain_spaces_and_alone():
    F.batch_of_plain_spaces_and_alone("cpu")


def test_every_joint_class_syncs_to_host():
    F.joints_sync_to_host("cpu")


def test_faults_report_like_a_debug_build():
    F.faults_report_like_a_debug_build("cpu")
