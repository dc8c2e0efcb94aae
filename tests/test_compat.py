"""An application written against the reference's module names runs
unchanged through polymage_amd/compat (CPU: build and plan; the GPU run of
the same library is in test_gpu_parity)."""
import os
import subprocess
import sys
import textwrap

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_style_app_builds(tmp_path):
    app = textwrap.dedent('''
        import sys
        from compiler import *
        from constructs import *
        from utils import get_ordered_cfunc_params
        R = Parameter(Int, "R"); C = Parameter(Int, "C")
        x = Variable(Int, "x"); y = Variable(Int, "y")
        img = Image(Float, "img", [R + 2, C + 2])
        row = Interval(Int, 0, R + 1); col = Interval(Int, 0, C + 1)
        cond = Condition(x, '>=', 1) & Condition(x, '<=', R) & Condition(y, '>=', 1) & Condition(y, '<=', C)
        blur = Function(([x, y], [row, col]), Float, "blur")
        blur.defn = [Case(cond, Stencil(img, [x, y], [[1, 2, 1], [2, 4, 2], [1, 2, 1]]) / 16.0)]
        pipe = buildPipeline([blur], param_estimates=[(R, 1000), (C, 800)],
                             param_constraints=[Condition(R, "==", 1000), Condition(C, "==", 800)],
                             tile_sizes=[32, 64], group_size=5, pipe_name="blur")
        params, inputs, outputs = get_ordered_cfunc_params(pipe)
        print([p.name for p in params], [i.name for i in inputs], [o.name for o in outputs])
        code = pipe.generate_code(is_extern_c_func=True, are_io_void_ptrs=True)
        src = str(code)
        assert 'extern "C" int pipeline_blur_dev' in src and 'void pipeline_blur(' in src
        pipe.pipeline_graph.write(sys.argv[1])
        print("ok")
    ''')
    f = tmp_path / "app.py"
    f.write_text(app)
    env = dict(os.environ, PYTHONPATH=os.pathsep.join([os.path.join(ROOT, 'polymage_amd', 'compat'), ROOT]))
    out = subprocess.check_output([sys.executable, str(f), str(tmp_path / 'g.dot')], env=env, cwd=str(tmp_path))
    assert b"['C', 'R'] ['img'] ['blur']" in out and out.strip().endswith(b"ok")
    assert 'digraph' in (tmp_path / 'g.dot').read_text()
