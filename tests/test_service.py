"""The service adapter's wire contract without a GPU: the settings and image
messages encode to the bytes protobuf gives for the field numbers and types
of pars/proto/settings.proto:5-27 and image.proto:5-9 (known answers written
out by hand from the protobuf encoding rules), and the PAIR loop of
service.cpp:37-67 answers each request in order."""
import struct

import pytest

from pars_amd import service as S


def _key(field, wire):
    return bytes([(field << 3) | wire])


def test_settings_encoding_known_answer():
    s = S.SETTINGS(mode="streamlines", dataset_filepath="a.h5", seed_generation_stride=[2, 3, 4],
                   seed_generation_iterations=300, particle_tracing_integrator="runge_kutta_4",
                   particle_tracing_step_size=0.5, particle_tracing_load_balance=True,
                   color_generation_mode="rgb", color_generation_free_parameter=0.25,
                   raytracing_camera_position=[1.0, 2.0, 3.0], raytracing_image_size=[640, 480],
                   raytracing_streamline_radius=0.125, raytracing_iterations=8)
    want = b"".join([
        _key(1, 2), bytes([11]), b"streamlines",
        _key(3, 2), bytes([4]), b"a.h5",
        _key(4, 2), bytes([3, 2, 3, 4]),                      # packed repeated int32
        _key(5, 0), bytes([0xAC, 0x02]),                       # varint 300
        _key(6, 2), bytes([13]), b"runge_kutta_4",
        _key(7, 5), struct.pack("<f", 0.5),
        _key(8, 0), bytes([1]),
        _key(9, 2), bytes([3]), b"rgb",
        _key(10, 5), struct.pack("<f", 0.25),
        _key(11, 2), bytes([12]), struct.pack("<3f", 1.0, 2.0, 3.0),
        _key(14, 2), bytes([4, 0x80, 0x05, 0xE0, 0x03]),       # 640, 480
        _key(15, 5), struct.pack("<f", 0.125),
        bytes([0x80, 0x01]), bytes([8]),                       # field 16, varint: a two-byte key
    ])
    assert s.SerializeToString() == want
    back = S.SETTINGS()
    back.ParseFromString(want)
    assert back == s


def test_image_encoding_known_answer():
    img = S.IMAGE(size=[2, 1], data=bytes(range(8)))
    assert img.SerializeToString() == _key(1, 2) + bytes([2, 2, 1]) + _key(2, 2) + bytes([8]) + bytes(range(8))


class _Echo:
    def __init__(self):
        self.seen = []

    def handle(self, request):
        s = S.SETTINGS()
        s.ParseFromString(request)
        self.seen.append(s.seed_generation_iterations)
        return S.IMAGE(size=[s.seed_generation_iterations, 1]).SerializeToString()


def test_pair_loop_answers_in_order():
    ch = S.PairChannel()
    for k in range(5):
        ch.client.send(S.SETTINGS(seed_generation_iterations=k + 1).SerializeToString())
    p = _Echo()
    assert S.run(ch.service, p, max_requests=5) == 5
    got = []
    for _ in range(5):
        img = S.IMAGE()
        img.ParseFromString(ch.client.recv(timeout=1))
        got.append(img.size[0])
    assert got == [1, 2, 3, 4, 5] and p.seen == [1, 2, 3, 4, 5]
