"""Transformation-tree semantics of the Computation API, without a GPU.

The cases follow reikna's own API tests (tests/test_core/test_transformation.py:
69-117 signature rewriting, :297-464 connection errors;
tests/test_core/test_computation.py:62-76 parameter names): the parameter order of
a signature after transformations are connected, and the ValueError texts of
every refused connection.  The computation below has reikna's test computation's
signature (tests/test_core/dummy.py:57-66: outputs C, D, inputs A, B, scalar
coeff); only its signature is used, so no plan is built.
"""

import re

import numpy
import pytest

from reikna_amd import functions
from reikna_amd.core import Annotation, Computation, Parameter, Transformation
from reikna_amd.runtime import ArrayMetadata


class DummyAdvanced(Computation):
    # the io-parameter computation of tests/test_core/dummy.py:253-262
    def __init__(self, arr, coeff):
        Computation.__init__(self, [
            Parameter("C", Annotation(arr, "io")),
            Parameter("D", Annotation(arr, "io")),
            Parameter("coeff1", Annotation(coeff)),
            Parameter("coeff2", Annotation(coeff)),
        ])


class Dummy(Computation):
    def __init__(self, arr1, arr2, coeff):
        Computation.__init__(self, [
            Parameter("C", Annotation(arr1, "o")),
            Parameter("D", Annotation(arr2, "o")),
            Parameter("A", Annotation(arr1, "i")),
            Parameter("B", Annotation(arr2, "i")),
            Parameter("coeff", Annotation(coeff)),
        ])


def tr_identity(arr):
    return Transformation(
        [Parameter("o1", Annotation(arr, "o")), Parameter("i1", Annotation(arr, "i"))],
        "${o1.store_same}(${i1.load_same});",
    )


def tr_2_to_1(arr, scalar):
    return Transformation(
        [
            Parameter("o1", Annotation(arr, "o")),
            Parameter("i1", Annotation(arr, "i")),
            Parameter("i2", Annotation(arr, "i")),
            Parameter("s1", Annotation(scalar)),
        ],
        "${o1.store_same}(${mul}(${i1.load_same}, ${s1}) + ${i2.load_same});",
        render_kwds=dict(mul=functions.mul(arr.dtype, scalar.dtype, out_dtype=arr.dtype)),
    )


def tr_1_to_2(arr):
    return Transformation(
        [
            Parameter("o1", Annotation(arr, "o")),
            Parameter("o2", Annotation(arr, "o")),
            Parameter("i1", Annotation(arr, "i")),
        ],
        "${o1.store_same}(${i1.load_same}); ${o2.store_same}(${i1.load_same});",
    )


def tr_scale(arr, coeff_dtype):
    return Transformation(
        [
            Parameter("o1", Annotation(arr, "o")),
            Parameter("i1", Annotation(arr, "i")),
            Parameter("s1", Annotation(coeff_dtype)),
        ],
        "${o1.store_same}(${mul}(${i1.load_same}, ${s1}));",
        render_kwds=dict(mul=functions.mul(arr.dtype, coeff_dtype, out_dtype=arr.dtype)),
    )


ARR = ArrayMetadata((200, 200), numpy.complex64)
COEFF = numpy.float32


def dummy():
    return Dummy(ARR, ARR, COEFF)


def test_io_parameter_in_transformation():
    with pytest.raises(ValueError, match=re.escape("Transformation cannot have 'io' parameters ('o1')")):
        Transformation(
            [Parameter("o1", Annotation(ArrayMetadata(100, numpy.float32), "io"))],
            "${o1.store_same}(${o1.load_same});",
        )


def test_signature_correctness():
    d = dummy()
    assert list(d.signature.parameters.values()) == [
        Parameter("C", Annotation(ARR, "o")),
        Parameter("D", Annotation(ARR, "o")),
        Parameter("A", Annotation(ARR, "i")),
        Parameter("B", Annotation(ARR, "i")),
        Parameter("coeff", Annotation(COEFF)),
    ]
    identity = tr_identity(d.parameter.A)
    join = tr_2_to_1(d.parameter.A, d.parameter.coeff)
    split = tr_1_to_2(d.parameter.A)
    scale = tr_scale(d.parameter.A, d.parameter.coeff.dtype)

    d.parameter.A.connect(identity, identity.o1, A_prime=identity.i1)
    d.parameter.B.connect(join, join.o1, A_prime=join.i1, B_prime=join.i2, B_param=join.s1)
    d.parameter.B_prime.connect(identity, identity.o1, B_new_prime=identity.i1)
    d.parameter.C.connect(split, split.i1, C_half1=split.o1, C_half2=split.o2)
    d.parameter.C_half1.connect(identity, identity.i1, C_new_half1=identity.o1)
    d.parameter.D.connect(scale, scale.i1, D_prime=scale.o1, D_param=scale.s1)

    # outputs first, then inputs, then scalars, each in depth-first tree order
    assert list(d.signature.parameters.values()) == [
        Parameter("C_new_half1", Annotation(ARR, "o")),
        Parameter("C_half2", Annotation(ARR, "o")),
        Parameter("D_prime", Annotation(ARR, "o")),
        Parameter("D_param", Annotation(COEFF)),
        Parameter("A_prime", Annotation(ARR, "i")),
        Parameter("B_new_prime", Annotation(ARR, "i")),
        Parameter("B_param", Annotation(COEFF)),
        Parameter("coeff", Annotation(COEFF)),
    ]


def test_connect_returns_the_computation():
    d = dummy()
    identity = tr_identity(d.parameter.A)
    assert d.connect("A", identity, identity.o1, A_prime=identity.i1) is d
    assert "A_prime" in d.signature.parameters


def test_alien_parameters():
    d, d2 = dummy(), dummy()
    identity = tr_identity(d.parameter.A)
    identity2 = tr_identity(d.parameter.A)
    with pytest.raises(ValueError, match=re.escape("The connection target must belong to this computation.")):
        d.connect(d2.parameter.A, identity, "o1", A_prime="i1")
    with pytest.raises(ValueError, match="The transformation parameter must belong to the provided transformation"):
        d.connect(d.parameter.A, identity, identity2.o1, A_prime="i1")
    with pytest.raises(ValueError, match="The transformation parameter must belong to the provided transformation"):
        d.parameter.A.connect(identity, identity.o1, A_prime=identity2.i1)


def test_connector_repetition():
    d = dummy()
    identity = tr_identity(d.parameter.A)
    with pytest.raises(
        ValueError,
        match="Parameter 'A' cannot be supplied both as the main connector and one of the child connections",
    ):
        d.parameter.A.connect(identity, identity.o1, A=identity.o1, A_prime=identity.i1)


def test_wrong_connector():
    d = dummy()
    identity = tr_identity(d.parameter.A)
    d.parameter.A.connect(identity, identity.o1, A_prime=identity.i1)
    with pytest.raises(ValueError, match="Parameter 'AA' is not a part of the signature"):
        d.connect("AA", identity, identity.o1, A_pp=identity.i1)
    # 'A' is now an internal node, hidden behind the identity transformation
    with pytest.raises(ValueError, match="Parameter 'A' is hidden by transformations"):
        d.connect("B", identity, identity.o1, A=identity.i1)


def test_type_mismatch():
    d = dummy()
    wrong = tr_identity(ArrayMetadata((200, 100), numpy.complex64))
    with pytest.raises(ValueError, match=r"Incompatible types of the transformation parameter 'o1' .+ and the node 'A'"):
        d.connect("A", wrong, wrong.o1, A_prime=wrong.i1)


def test_wrong_data_path():
    d = DummyAdvanced(ARR, COEFF)
    identity = tr_identity(d.parameter.C)
    d.parameter.C.connect(identity, identity.o1, C_in=identity.i1)
    d.parameter.D.connect(identity, identity.i1, D_out=identity.o1)
    assert list(d.signature.parameters.values()) == [
        Parameter("C", Annotation(ARR, "o")),
        Parameter("C_in", Annotation(ARR, "i")),
        Parameter("D_out", Annotation(ARR, "o")),
        Parameter("D", Annotation(ARR, "i")),
        Parameter("coeff1", Annotation(COEFF)),
        Parameter("coeff2", Annotation(COEFF)),
    ]
    # the input side of C is hidden by the first connection, the output side of D by the second
    with pytest.raises(ValueError, match="'C' is not an input node"):
        d.parameter.C.connect(identity, identity.o1, C_in_prime=identity.i1)
    with pytest.raises(ValueError, match="'D' is not an output node"):
        d.parameter.D.connect(identity, identity.i1, D_out_prime=identity.o1)
    with pytest.raises(
        ValueError, match="Cannot connect transformation parameter 'o1' to an existing output node 'D_out'"
    ):
        d.parameter.C.connect(identity, identity.i1, D_out=identity.o1)
    # the output side of C is still free
    d.parameter.C.connect(identity, identity.i1, C_out=identity.o1)
    assert list(d.signature.parameters.values()) == [
        Parameter("C_out", Annotation(ARR, "o")),
        Parameter("C_in", Annotation(ARR, "i")),
        Parameter("D_out", Annotation(ARR, "o")),
        Parameter("D", Annotation(ARR, "i")),
        Parameter("coeff1", Annotation(COEFF)),
        Parameter("coeff2", Annotation(COEFF)),
    ]


def test_wrong_transformation_parameters():
    d = dummy()
    identity = tr_identity(d.parameter.A)
    with pytest.raises(ValueError, match=re.escape(
            "Supplied transformation names ['input', 'o1'] do not fully coincide with the existing ones: ['i1', 'o1']")):
        d.parameter.A.connect(identity, identity.o1, A_prime="input")
    with pytest.raises(ValueError, match=re.escape(
            "Supplied transformation names ['i1', 'i2', 'o1'] "
            "do not fully coincide with the existing ones: ['i1', 'o1']")):
        d.parameter.A.connect(identity, identity.o1, A_prime=identity.i1, A2="i2")


def test_underscore_names_refused():
    # tests/test_core/test_computation.py:62-76
    with pytest.raises(ValueError, match=re.escape("External parameter name cannot start with the underscore.")):
        Computation([Parameter("_C", Annotation(ARR, "o"))])
